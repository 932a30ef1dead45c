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
#include "scan_common.h"

// Tiles of blockDim markers per marker block (location, scatter, advection):
// each block stages its instance's grid in LDS once for all of them (4 tiles:
// location 0.86 -> 0.68 ms, advection 3.17 -> 3.04 ms on the 64-instance step).
#define ENS_MTILES 4

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

// Groups of one (method, field count) share a launch: launch row blockIdx.y
// is group gi[blockIdx.y]; a row's blocks beyond that group's count exit.
#define KB_MAX_MULTI 8
struct KbGroups {
  int gi[KB_MAX_MULTI], sx[KB_MAX_MULTI], sy[KB_MAX_MULTI];
};

enum { PS_ZERO, PS_LOCATE, PS_SCAN, PS_SCATTER, PS_ORDER, PS_PART, PS_NODES, PS_SUBT, PS_NODES2, PS_ADVECT, PS_N };
#ifndef PS_CHUNK
#define PS_CHUNK 32
#endif
#define PS_SPIN_MAX (1 << 22)
#define PS_PEND 64
#define PS_SAT 16
#define PS_ZERO_ITEM 4096

struct PersistArgs {
  const InstDev *insts;
  int n, ng, nrow, ncol;
  int64_t off[PS_N + 1];  // first item of each phase; off[PS_N] = total
  // rows: markers (256 per block), nodes, dTsub lattice, universal lattice
  const int64_t *mblk, *nblk, *sblk, *ublk;
  const int32_t *mbi, *nbi, *sbi, *ubi;
  const int64_t *cell_off, *tile_off;
  const int32_t *tile_inst;
  const int32_t *scan_upto;           // [scan tile]: its counts are complete when instances 0..upto are located
  const int32_t *scan_lo, *scan_hi;   // [scan tile]: instances whose scan the tile completes
  int32_t *done;                      // [PS_N][n] items credited
  const int32_t *need;                // [PS_N][n] items per (phase, instance)
  unsigned long long *ticket;         // [0] next chunk, [1] workgroups finished
  const int64_t *chunk;               // [nchunk + 1] first item of each chunk
  int64_t nchunk;
  int32_t *key, *slot, *perm, *count, *start, *counters, *scan_fail;
  unsigned long long *scan_state;
  int64_t nbucket, nstate, nmark;
  double *xs, *ys, *pcols, *part, *dTsub, *deltaT;
  const double *mt;  // payload column of marker T
  int32_t *ms;
  const double *const *colsrc;
  const ProjGroup *groups_sorted, *groups;
  const int64_t *partoff;
  const NodeGroup *nodes;
  KbGroups kg;
};

struct fvm_ensemble {
  int n = 0, ng = 0;
  fvm_tuning tune = {256, 8, 0};
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
  // persistent step (tune.persistent): work-item schedule and completion state
  int persist_nf = 0;                  // stored fields of the single partials kind, 0: not possible
  PersistArgs pa;                      // launch arguments, rebuilt with the block rows
  PersistArgs *d_pa = nullptr;         // their device copy (the kernel reads them through a pointer)
  int persist_grid = 0;                // resident workgroups
  int32_t *d_done = nullptr, *d_need = nullptr, *d_scan_upto = nullptr, *d_scan_lo = nullptr,
          *d_scan_hi = nullptr;
  unsigned long long *d_ticket = nullptr;
  size_t scan_cap = 0;
  int64_t *d_chunk = nullptr;
  size_t chunk_cap = 0;
  std::vector<int64_t> hcell;          // [n + 1] first basic cell, then [n + 1] first 8 x 8 tile, per instance
};

// ---- block -> instance ----------------------------------------------------------------
// Every kind of work pads each instance's range to whole blocks; a table maps
// a block (row table `bi`) or an ordering tile to its instance, so a block
// finds its instance with one load.

// ---- per-block bodies ---------------------------------------------------------------------
// Each body does the work of one block of one phase for instance I; `lb` is
// the block's index within that instance's range.  The per-phase kernels map
// blockIdx.x to (instance, lb); the persistent step (kb_persist) maps its work
// items the same way, so both run the same code.

// x[0..NX-1], y[0..NY-1], xc[0..NX], yc[0..NY] of the instance in LDS
__device__ __forceinline__ DGrid ens_stage_grid(const DGrid &g, double *sgrid) {
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
  return gl;
}

// gl: the instance's grid as ens_stage_grid left it in LDS
__device__ __forceinline__ void ens_locate_block(const InstDev &I, int64_t lb, const DGrid &gl, int32_t *__restrict__ key,
                                                 int32_t *__restrict__ slot, int32_t *__restrict__ count,
                                                 int32_t *__restrict__ counters) {
  const DGrid &g = gl;
  const double *gx = g.x, *gy = g.y, *gxc = g.xc, *gyc = g.yc;
  const int64_t m = lb * (int64_t)blockDim.x + threadIdx.x;
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

__device__ __forceinline__ void ens_scatter_block(const InstDev &I, int64_t lb, const int32_t *__restrict__ key,
                                                  const int32_t *__restrict__ slot, const int32_t *__restrict__ start,
                                                  int32_t *__restrict__ perm) {
  const int64_t m = lb * (int64_t)blockDim.x + threadIdx.x;
  if (m >= I.nMark) return;
  const int k = key[I.mark_off + m];
  if (k < 0) return;
  perm[start[k] + slot[I.mark_off + m]] = (int32_t)m;
}

// Bucket order of two basic cells of one instance (c1 < 0: none) by the
// calling wavefront (bucket_order_pair), with the payload gathered from the
// instance's columns.
__device__ __forceinline__ void ens_order_pair(const InstDev &I, int ii, int n, int64_t c0, int64_t c1,
                                               const int32_t *__restrict__ start, int32_t *__restrict__ perm,
                                               double *__restrict__ xs, double *__restrict__ ys,
                                               int32_t *__restrict__ ms, int ncol,
                                               const double *const *__restrict__ colsrc, double *__restrict__ pcols,
                                               int64_t nmark) {
  const double *__restrict__ X = I.X, *__restrict__ Y = I.Y;
  const int32_t *__restrict__ Mat = I.Mat;
  bucket_order_pair(c0, c1, start, perm, [&](int p, int v) {
    xs[p] = X[v];
    ys[p] = Y[v];
    ms[p] = Mat[v];
    for (int k = 0; k < ncol; k++) pcols[(int64_t)k * nmark + p] = colsrc[(int64_t)k * n + ii][v];
  });
}

// One block of the universal lattice ((NX+1) x (NY+1) cell coordinates) of
// instance ii for projection group gi.
template <int METHOD, int NF, int G>
__device__ __forceinline__ void ens_partials_block(const InstDev &I, int ii, int n, int64_t lb, int gi, int sx, int sy,
                                                   const ProjGroup *__restrict__ all_groups,
                                                   const int64_t *__restrict__ all_partoff,
                                                   const int32_t *__restrict__ start, const double *__restrict__ xs,
                                                   const double *__restrict__ ys, const int32_t *__restrict__ ms,
                                                   double *__restrict__ part) {
  const int64_t lt = lb * (int64_t)blockDim.x + threadIdx.x;
  const int64_t t = lt / G;
  const int j = (int)(lt % G);
  const int64_t nT = (int64_t)(I.g.NX + 1) * (I.g.NY + 1);
  const bool inside = t < nT;
  // 32-bit division (an instance's lattice has far fewer than 2^31 cells)
  const uint32_t t32 = (uint32_t)t, w32 = (uint32_t)(I.g.NX + 1);
  const int cx = inside ? (int)(t32 % w32) - 1 : -2;
  const int cy = inside ? (int)(t32 / w32) - 1 : -2;
  const Lat Lx = lattice(sx, I.g.NX), Ly = lattice(sy, I.g.NY);
  const bool live = cx >= Lx.lo && cx < Lx.lo + Lx.n && cy >= Ly.lo && cy < Ly.lo + Ly.n;
  const int64_t nL = (int64_t)Lx.n * Ly.n;
  const int64_t tl = live ? (int64_t)(cy - Ly.lo) * Lx.n + (cx - Lx.lo) : 0;
  const ProjGroup &pg = all_groups[(int64_t)gi * n + ii];
  __shared__ double smat[NF * FVM_MAXMAT];
  const double *mtab = stage_mat_tables<NF>(pg, smat);
  partials_cell_sorted<METHOD, NF, G>(I.g, pg, start + I.bucket_off, xs, ys, ms, mtab,
                                      part + all_partoff[(int64_t)gi * n + ii], nL, tl, live ? cx : 0,
                                      live ? cy : 0, j, live);
}

__device__ __forceinline__ void ens_nodes_block(const InstDev &I, int64_t lb, const NodeGroup &q) {
  const int64_t i = lb * (int64_t)blockDim.x + threadIdx.x;
  if (i >= (int64_t)I.g.NX * I.g.NY) return;
  node_group(q, (int)(i % I.g.NX), (int)(i / I.g.NX), i);
}

template <int G>
__device__ __forceinline__ void ens_subt_block(const InstDev &I, int ii, int64_t lb,
                                               const ProjGroup *__restrict__ groups,
                                               const int64_t *__restrict__ partoff, const int32_t *__restrict__ start,
                                               const int32_t *__restrict__ perm, const double *__restrict__ xs,
                                               const double *__restrict__ ys, const double *__restrict__ mt,
                                               double *__restrict__ part, double *__restrict__ dTsub) {
  const int nLx = I.g.NX - 1, nLy = I.g.NY - 1;
  const int64_t nL = (int64_t)nLx * nLy;
  const int64_t lt = lb * (int64_t)blockDim.x + threadIdx.x;
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

// gl: the instance's grid staged in LDS (ens_stage_grid), as in the location pass
__device__ __forceinline__ void ens_advect_block(const InstDev &I, int64_t lb, const DGrid &gl,
                                                 const double *__restrict__ deltaT, const double *__restrict__ dTsub) {
  const int64_t m = lb * (int64_t)blockDim.x + threadIdx.x;
  if (m >= I.nMark) return;
  advect_marker<false>(m, gl, I.X, I.Y, I.VX, I.VY, I.wxy, I.cellX, I.cellY, I.vx, I.vy, I.wxyn, I.bv, I.dt, I.T,
                       I.Tdot, deltaT + I.node_off, dTsub + I.mark_off, I.dt);
  // the rotation fused behind the RK4 stages: the batched launch keeps one kernel per phase
  rotate_eii_marker(m, I.g.ncx(), I.g.NY, I.cellX, I.cellY, I.wxy, I.sxx, I.syy, I.sxy, I.exx, I.eyy, I.exy, I.Eii,
                    I.dt, I.jaumann);
}

// ---- per-phase kernels ------------------------------------------------------------------
// Every kind of work pads each instance's range to whole blocks; a table maps
// a block (row table `bi`) or an ordering tile to its instance, so a block
// finds its instance with one load.
__global__ void __launch_bounds__(512) kb_locate(const InstDev *__restrict__ insts, int n,
                                                 const int64_t *__restrict__ blk, const int32_t *__restrict__ bi,
                                                 int32_t *__restrict__ key,
                                                 int32_t *__restrict__ slot, int32_t *__restrict__ count,
                                                 int32_t *__restrict__ counters) {
  extern __shared__ double sgrid[];
  const int ii = bi[blockIdx.x];
  const DGrid gl = ens_stage_grid(insts[ii].g, sgrid);
  for (int t = 0; t < ENS_MTILES; t++)
    ens_locate_block(insts[ii], (blockIdx.x - blk[ii]) * ENS_MTILES + t, gl, key, slot, count, counters);
}

__global__ void __launch_bounds__(512) kb_scatter(const InstDev *__restrict__ insts, int n,
                                                  const int64_t *__restrict__ blk, const int32_t *__restrict__ bi,
                                                  const int32_t *__restrict__ key, const int32_t *__restrict__ slot,
                                                  const int32_t *__restrict__ start, int32_t *__restrict__ perm) {
  const int ii = bi[blockIdx.x];
  for (int t = 0; t < ENS_MTILES; t++)
    ens_scatter_block(insts[ii], (blockIdx.x - blk[ii]) * ENS_MTILES + t, key, slot, start, perm);
}

// Bucket order of every instance's cells (two cells per wavefront, one above
// the other, bucket_order_pair) with the payload gathered from the owning
// instance's columns.  Cells are visited in
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
  const int64_t tile = w >> 5;
  if (tile >= tile_off[n]) return;  // whole wavefront exits together
  const int ii = tile_inst[tile];
  const InstDev &I = insts[ii];
  const int ncx = I.g.NX - 1, ncy = I.g.NY - 1;
  const uint32_t ntx = (uint32_t)(ncx + 7) >> 3;
  const uint32_t tl = (uint32_t)(tile - tile_off[ii]);
  // cells l and l + 1 of the tile: one above the other
  const int l = (int)(w & 31) * 2;
  const int cx = (int)(tl % ntx) * 8 + (l >> 3), cy = (int)(tl / ntx) * 8 + (l & 7);
  if (cx >= ncx || cy >= ncy) return;
  const int64_t c0 = cell_off[ii] + (int64_t)cy * ncx + cx;
  ens_order_pair(I, ii, n, c0, cy + 1 < ncy ? c0 + ncx : -1, start, perm, xs, ys, ms, ncol, colsrc, pcols, nmark);
}

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
  const int ii = ubi[ub];
  ens_partials_block<METHOD, NF, G>(insts[ii], ii, n, ub - ublk[ii], kg.gi[q], kg.sx[q], kg.sy[q], all_groups,
                                    all_partoff, start, xs, ys, ms, part);
}

// node pass: launch row blockIdx.y = group g0 + blockIdx.y, every field of it
__global__ void __launch_bounds__(256) kb_nodes(const InstDev *__restrict__ insts, int n,
                                                const int64_t *__restrict__ blk, const int32_t *__restrict__ bi,
                                                const NodeGroup *__restrict__ nodes, int g0) {
  const int ii = bi[blockIdx.x];
  ens_nodes_block(insts[ii], blockIdx.x - blk[ii], nodes[(int64_t)(g0 + blockIdx.y) * n + ii]);
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
  ens_subt_block<G>(insts[ii], ii, blockIdx.x - blk[ii], groups, partoff, start, perm, xs, ys, mt, part, dTsub);
}

// five waves per SIMD: the tile loop took the body from 90 to 100 VGPRs (four
// waves); held to 94 without spilling, the phase runs 2.95 ms instead of 3.06
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(5))) kb_advect(const InstDev *__restrict__ insts, int n,
                                                 const int64_t *__restrict__ blk, const int32_t *__restrict__ bi,
                                                 const double *__restrict__ deltaT, const double *__restrict__ dTsub) {
  extern __shared__ double sgrid[];
  const int ii = bi[blockIdx.x];
  const DGrid gl = ens_stage_grid(insts[ii].g, sgrid);
  for (int t = 0; t < ENS_MTILES; t++)
    ens_advect_block(insts[ii], (blockIdx.x - blk[ii]) * ENS_MTILES + t, gl, deltaT, dTsub);
}

// ---- persistent step ------------------------------------------------------------------------
// The whole step as ONE launch: a grid of resident workgroups takes chunks of
// consecutive work items from a ticket and runs each item's body (the same
// bodies as the per-phase kernels).  A chunk is PS_CHUNK items, an eighth of
// that for the ordering pass's 64-cell tiles, and one tile of the scan: its
// tiles chain through the look-back, so consecutive tiles must run side by
// side, not one after another in one workgroup.  Items are numbered phase by phase in
// step order (zero, locate, scan, scatter, order, partials, nodes, dTsub
// partials, dTsub nodes, advect), instances in order within a phase, so every
// item depends only on items numbered before it: an item waits until the
// previous phase of its own instance is complete (the scan: until the counts
// of every bucket up to its tile are), and instances overlap freely across
// phase boundaries.
//
// Completion is counted per (phase, instance).  A workgroup credits its items
// when its chunk ends (one release per chunk) and counts its own finished but
// uncredited items as done when it checks a dependency, so the lowest
// uncredited item can always run: no wait can cycle, and every wait is also
// bounded (PS_SPIN_MAX polls; on expiry the step stops and sets a sticky flag,
// counters[2], which fvm_ensemble_counters reports).  The last workgroup to
// leave resets the ticket and the completion counts for the next launch.
struct PsPend {
  int n;
  int ph[PS_PEND], inst[PS_PEND], cnt[PS_PEND];
  // (phase, instance) pairs this workgroup has seen complete and acquired:
  // their data is final, so later items that need them wait and fence no more
  int nsat;
  int sat[PS_SAT];
};

__device__ __forceinline__ int ps_pending(const PsPend &p, int ph, int i) {
  int c = 0;
  for (int k = 0; k < p.n; k++) c += (p.ph[k] == ph && p.inst[k] == i) ? p.cnt[k] : 0;
  return c;
}

// thread 0: credit the pending items (after a release, so that everything the
// workgroup wrote is visible before any count that vouches for it)
__device__ __forceinline__ void ps_flush(const PersistArgs &A, PsPend &p) {
  if (p.n == 0) return;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  for (int k = 0; k < p.n; k++)
    __hip_atomic_fetch_add(&A.done[p.ph[k] * A.n + p.inst[k]], p.cnt[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  p.n = 0;
}

__device__ __forceinline__ void ps_credit(const PersistArgs &A, PsPend &p, int ph, int lo, int hi, int cnt) {
  for (int i = lo; i <= hi; i++) {
    int k = 0;
    while (k < p.n && !(p.ph[k] == ph && p.inst[k] == i)) k++;
    if (k == p.n) {
      if (p.n == PS_PEND) {
        ps_flush(A, p);
        k = 0;
      }
      p.ph[k] = ph, p.inst[k] = i, p.cnt[k] = 0;
      p.n = k + 1 > p.n ? k + 1 : p.n;
    }
    p.cnt[k] += cnt;
  }
}

// thread 0: wait until phase `ph` is complete for instances lo..hi; false if
// the wait ran out or another workgroup gave up
__device__ __forceinline__ bool ps_wait(const PersistArgs &A, PsPend &p, int ph, int lo, int hi) {
  bool global = false;
  for (int i = lo; i <= hi; i++) {
    const int key = ph * A.n + i;
    bool seen = false;
    for (int k = 0; k < p.nsat; k++) seen |= p.sat[k] == key;
    if (seen) continue;
    const int want = A.need[ph * A.n + i] - ps_pending(p, ph, i);
    if (want <= 0) continue;
    global = true;
    int spins = 0;
    while (__hip_atomic_load(&A.done[ph * A.n + i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < want) {
      if (++spins > PS_SPIN_MAX || __hip_atomic_load(&A.counters[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
        atomicOr(&A.counters[2], 1);
        return false;
      }
      __builtin_amdgcn_s_sleep(4);
    }
    // complete without this workgroup's pending items: remember it (the
    // acquire below covers it)
    p.sat[p.nsat < PS_SAT ? p.nsat : PS_SAT - 1] = key;
    if (p.nsat < PS_SAT) p.nsat++;
  }
  if (global) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  return true;
}

// Three waves per SIMD (168 VGPRs, a few spills) ran the 64-instance step in
// 18.6 ms against 20.5 ms at two (213 VGPRs, none); chunks of 8, 32 or 128
// items were within 10% of each other.
#ifndef PS_WAVES_PER_EU
#define PS_WAVES_PER_EU 3
#endif
// Runs of items: consecutive items of one phase and one instance share one
// dependency wait, one staging of the instance's grid and one credit, so the
// per-item cost is the body alone.
template <int NF, int G>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(PS_WAVES_PER_EU)))
kb_persist(const PersistArgs *__restrict__ Ap) {
  const PersistArgs &A = *Ap;
  extern __shared__ double sgrid[];
  __shared__ int64_t s_item;
  __shared__ int s_go;
  __shared__ PsPend pend;
  if (threadIdx.x == 0) pend.n = 0, pend.nsat = 0;
  bool stop = false;
  while (!stop) {
    if (threadIdx.x == 0) {
      const bool ab = __hip_atomic_load(&A.counters[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
      const int64_t c = ab ? A.nchunk : (int64_t)atomicAdd(&A.ticket[0], 1ull);
      s_item = c < A.nchunk ? c : A.nchunk;
    }
    __syncthreads();
    // the chunk number is the same in every lane: say so, so that everything
    // derived from it (instance record, rows, column pointers) is scalar
    const int64_t ch = (int64_t)(((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(s_item >> 32)) << 32) |
                                 (uint32_t)__builtin_amdgcn_readfirstlane((int)s_item));
    if (ch >= A.nchunk) break;
    const int64_t it1 = A.chunk[ch + 1];
    int ph = 0;
    for (int64_t it = A.chunk[ch]; it < it1;) {
      while (it >= A.off[ph + 1]) ph++;
      const int64_t b = it - A.off[ph];
      // the run's instance, its end (the instance's last item of this phase, or
      // the chunk's end), dependency (phase dph for instances dlo..dhi) and
      // credit range
      int ii = 0, dph = -1, dlo = 0, dhi = 0, clo = 0, chi = 0;
      int64_t lb = 0, bend = A.off[ph + 1] - A.off[ph];
      switch (ph) {
        case PS_ZERO: break;
        case PS_LOCATE: case PS_SCATTER: case PS_ADVECT:
          ii = A.mbi[b], lb = b - A.mblk[ii], bend = A.mblk[ii + 1];
          break;
        case PS_SCAN:
          dph = PS_LOCATE, dlo = 0, dhi = A.scan_upto[b], clo = A.scan_lo[b], chi = A.scan_hi[b], bend = b + 1;
          break;
        case PS_ORDER: ii = A.tile_inst[b], lb = b - A.tile_off[ii], bend = A.tile_off[ii + 1]; break;
        case PS_PART: ii = A.ubi[b], lb = b - A.ublk[ii], bend = A.ublk[ii + 1]; break;
        case PS_NODES: case PS_NODES2: ii = A.nbi[b], lb = b - A.nblk[ii], bend = A.nblk[ii + 1]; break;
        case PS_SUBT: ii = A.sbi[b], lb = b - A.sblk[ii], bend = A.sblk[ii + 1]; break;
      }
      if (ph == PS_LOCATE) dph = PS_ZERO, dlo = dhi = 0;
      else if (ph > PS_SCAN) dph = ph - 1, dlo = dhi = ii;
      if (ph > PS_SCAN || ph == PS_LOCATE) clo = chi = ii;
      const int64_t rend = A.off[ph] + bend < it1 ? A.off[ph] + bend : it1;
      const int nrun = (int)(rend - it);
      if (threadIdx.x == 0) s_go = dph < 0 ? 1 : (ps_wait(A, pend, dph, dlo, dhi) ? 1 : 0);
      __syncthreads();
      if (!__builtin_amdgcn_readfirstlane(s_go)) {
        stop = true;
        break;
      }
      const InstDev &I = A.insts[ii];
      switch (ph) {
        case PS_ZERO:
          for (int64_t z = b; z < b + nrun; z++) {
            const int64_t z0 = z * PS_ZERO_ITEM;
            for (int64_t k = z0 + threadIdx.x; k < z0 + PS_ZERO_ITEM; k += blockDim.x) {
              if (k <= A.nbucket) A.count[k] = 0;
              if (k < A.nstate) A.scan_state[k] = 0ull;
            }
            if (z == 0 && threadIdx.x < 2) A.counters[threadIdx.x] = 0;
          }
          break;
        case PS_LOCATE: {
          const DGrid gl = ens_stage_grid(I.g, sgrid);
          for (int r = 0; r < nrun; r++)
            for (int t = 0; t < ENS_MTILES; t++)
              ens_locate_block(I, (lb + r) * ENS_MTILES + t, gl, A.key, A.slot, A.count, A.counters);
          break;
        }
        case PS_SCAN: scan_tile((int32_t)b, A.count, A.start, A.scan_state, A.nbucket, A.scan_fail); break;
        case PS_SCATTER:
          for (int r = 0; r < nrun; r++)
            for (int t = 0; t < ENS_MTILES; t++)
              ens_scatter_block(I, (lb + r) * ENS_MTILES + t, A.key, A.slot, A.start, A.perm);
          break;
        case PS_ORDER: {
          // each tile's 64 cells, 16 per wavefront in pairs (cell l at (l >> 3, l & 7)
          // in the tile, with cell l + 1 above it)
          const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
          const int ncx = I.g.NX - 1, ncy = I.g.NY - 1;
          const uint32_t ntx = (uint32_t)(ncx + 7) >> 3;
          for (int r = 0; r < nrun; r++) {
            const uint32_t t = (uint32_t)(lb + r);
            for (int l = 2 * wave; l < 64; l += 8) {
              const int cx = (int)(t % ntx) * 8 + (l >> 3), cy = (int)(t / ntx) * 8 + (l & 7);
              if (cx < ncx && cy < ncy) {
                const int64_t c0 = A.cell_off[ii] + (int64_t)cy * ncx + cx;
                ens_order_pair(I, ii, A.n, c0, cy + 1 < ncy ? c0 + ncx : -1, A.start, A.perm, A.xs, A.ys, A.ms,
                               A.ncol, A.colsrc, A.pcols, A.nmark);
              }
            }
          }
          break;
        }
        case PS_PART:
          for (int r = 0; r < nrun; r++)
            for (int q = 0; q < A.nrow; q++)
              ens_partials_block<PM_ARITH, NF, G>(I, ii, A.n, lb + r, A.kg.gi[q], A.kg.sx[q], A.kg.sy[q],
                                                  A.groups_sorted, A.partoff, A.start, A.xs, A.ys, A.ms, A.part);
          break;
        case PS_NODES:
          for (int r = 0; r < nrun; r++)
            for (int g = 0; g < A.ng; g++) ens_nodes_block(I, lb + r, A.nodes[(int64_t)g * A.n + ii]);
          break;
        case PS_SUBT:
          for (int r = 0; r < nrun; r++)
            ens_subt_block<G>(I, ii, lb + r, A.groups + (int64_t)A.ng * A.n, A.partoff + (int64_t)A.ng * A.n,
                              A.start, A.perm, A.xs, A.ys, A.mt, A.part, A.dTsub);
          break;
        case PS_NODES2:
          for (int r = 0; r < nrun; r++) ens_nodes_block(I, lb + r, A.nodes[(int64_t)A.ng * A.n + ii]);
          break;
        case PS_ADVECT: {
          const DGrid gl = ens_stage_grid(I.g, sgrid);
          for (int r = 0; r < nrun; r++)
            for (int t = 0; t < ENS_MTILES; t++) ens_advect_block(I, (lb + r) * ENS_MTILES + t, gl, A.deltaT, A.dTsub);
          break;
        }
      }
      __syncthreads();  // the run's writes (and its use of the staged grid) are done
      if (threadIdx.x == 0) ps_credit(A, pend, ph, clo, chi, nrun);
      it = rend;
    }
    if (threadIdx.x == 0) ps_flush(A, pend);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    ps_flush(A, pend);
    // acquire-release: the last one out sees every other workgroup's credits
    // before it clears them
    const unsigned long long prev =
        __hip_atomic_fetch_add(&A.ticket[1], 1ull, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    if (prev == (unsigned long long)gridDim.x - 1) {  // last one out: reset for the next launch
      for (int k = 0; k < PS_N * A.n; k++) A.done[k] = 0;
      A.ticket[0] = 0ull;
      A.ticket[1] = 0ull;
    }
  }
}

// ---- host side ----------------------------------------------------------------------------
static void free_dev(void *p) {
  if (p) (void)hipFree(p);
}

// block offsets rows: 0 markers (ENS_MTILES tiles of tune.block), 1 nodes (256), 2 + g lattice cells of
// group g, then the universal lattice ((NX+1) x (NY+1) cells) of the shared
// partials launches
static int64_t *blk_row(fvm_ensemble *E, int r) { return E->dblk + (int64_t)r * (E->n + 1); }
static const int32_t *bi_row(fvm_ensemble *E, int r) { return E->dbi + E->hbi_off[r]; }

static int urow(const fvm_ensemble *E) { return 2 + E->ng + 1; }

static int persist_setup(fvm_ctx *ctx, fvm_ensemble *E);

static int build_blocks(fvm_ctx *ctx, fvm_ensemble *E) {
  const int n = E->n, rows = urow(E) + 1;
  E->hblk.assign((size_t)rows * (n + 1), 0);
  const int B = E->tune.block, G = E->tune.lanes;
  for (int i = 0; i < n; i++) {
    const InstDev &I = E->hinst[i];
    const int64_t nm = I.nMark, nn = (int64_t)I.g.NX * I.g.NY;
    E->hblk[0 * (n + 1) + i + 1] = E->hblk[0 * (n + 1) + i] + (nm + (int64_t)B * ENS_MTILES - 1) / ((int64_t)B * ENS_MTILES);
    E->hblk[1 * (n + 1) + i + 1] = E->hblk[1 * (n + 1) + i] + (nn + 255) / 256;
    for (int gi = 0; gi <= E->ng; gi++) {
      const Lat Lx = lattice(E->gstag_x[gi], I.g.NX), Ly = lattice(E->gstag_y[gi], I.g.NY);
      // row 2 + ng: the dTsub pass, which takes at least 8 lanes per cell
      const int64_t th = (int64_t)Lx.n * Ly.n * (gi == E->ng && G < 8 ? 8 : G);
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
  return E->tune.persistent ? persist_setup(ctx, E) : FVM_OK;
}

// ---- persistent schedule (host) -------------------------------------------------------------
#define PS_DISPATCH(NF_, G_, CALL)                                 \
  switch ((NF_) * 100 + (G_)) {                                    \
    case 108: CALL(1, 8); break;                                   \
    case 116: CALL(1, 16); break;                                  \
    case 132: CALL(1, 32); break;                                  \
    case 208: CALL(2, 8); break;                                   \
    case 216: CALL(2, 16); break;                                  \
    case 232: CALL(2, 32); break;                                  \
    case 308: CALL(3, 8); break;                                   \
    case 316: CALL(3, 16); break;                                  \
    case 332: CALL(3, 32); break;                                  \
    case 408: CALL(4, 8); break;                                   \
    case 416: CALL(4, 16); break;                                  \
    case 432: CALL(4, 32); break;                                  \
    case 508: CALL(5, 8); break;                                   \
    case 516: CALL(5, 16); break;                                  \
    case 532: CALL(5, 32); break;                                  \
    default: return fvm_set_err(ctx, FVM_ERR_ARG, "persistent step: no kernel for %d fields, %d lanes", NF_, G_); \
  }

// Item ranges of every phase, the completion counts each (phase, instance)
// needs, the scan tiles' instance ranges and the launch arguments; the
// completion counts and the ticket start at zero (every launch leaves them so).
static int persist_setup(fvm_ctx *ctx, fvm_ensemble *E) {
  const int n = E->n, u = urow(E);
  auto rowcnt = [&](int r, int i) { return E->hblk[(size_t)r * (n + 1) + i + 1] - E->hblk[(size_t)r * (n + 1) + i]; };
  auto rowtot = [&](int r) { return E->hblk[(size_t)r * (n + 1) + n]; };
  const int64_t T = SCAN_BS * SCAN_IPT;
  const int64_t ntiles = fvm_scan_tiles(E->nbucket), nstate = ntiles + 2;
  const int64_t nzero = (std::max(E->nbucket + 1, nstate) + PS_ZERO_ITEM - 1) / PS_ZERO_ITEM;
  std::vector<int32_t> need((size_t)PS_N * n, 0), upto((size_t)ntiles), lo((size_t)ntiles, n), hi((size_t)ntiles, -1);
  auto at = [&](int ph, int i) -> int32_t & { return need[(size_t)ph * n + i]; };
  at(PS_ZERO, 0) = (int32_t)nzero;
  for (int i = 0; i < n; i++) {
    const InstDev &I = E->hinst[i];
    at(PS_LOCATE, i) = at(PS_SCATTER, i) = at(PS_ADVECT, i) = (int32_t)rowcnt(0, i);
    at(PS_ORDER, i) = (int32_t)(E->hcell[n + 1 + i + 1] - E->hcell[n + 1 + i]);
    at(PS_PART, i) = (int32_t)rowcnt(u, i);
    at(PS_NODES, i) = at(PS_NODES2, i) = (int32_t)rowcnt(1, i);
    at(PS_SUBT, i) = (int32_t)rowcnt(2 + E->ng, i);
    // scan tiles whose output the instance reads: its buckets and the start of
    // the next instance's (or the total, written by the last tile)
    const int64_t b0 = I.bucket_off, b1 = I.bucket_off + 4LL * (I.g.NX - 1) * (I.g.NY - 1);
    const int64_t t0 = b0 / T, t1 = std::min(b1, E->nbucket - 1) / T;
    at(PS_SCAN, i) = (int32_t)(t1 - t0 + 1);
    for (int64_t t = t0; t <= t1; t++) {
      lo[(size_t)t] = std::min(lo[(size_t)t], i);
      hi[(size_t)t] = std::max(hi[(size_t)t], i);
    }
  }
  // tile t sums the counts of every bucket before its end: instances whose
  // buckets start before it
  for (int64_t t = 0, i = 0; t < ntiles; t++) {
    const int64_t end = std::min((t + 1) * T, E->nbucket);
    while (i + 1 < n && E->hinst[i + 1].bucket_off < end) i++;
    upto[(size_t)t] = (int32_t)i;
  }
  PersistArgs &A = E->pa;
  memset(&A, 0, sizeof(A));
  const int64_t cnt[PS_N] = {nzero, rowtot(0), ntiles, rowtot(0), E->ntile, rowtot(u), rowtot(1), rowtot(2 + E->ng),
                             rowtot(1), rowtot(0)};
  A.off[0] = 0;
  for (int ph = 0; ph < PS_N; ph++) A.off[ph + 1] = A.off[ph] + cnt[ph];
  if ((size_t)ntiles > E->scan_cap) {
    FVM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    free_dev(E->d_scan_upto), free_dev(E->d_scan_lo), free_dev(E->d_scan_hi);
    E->d_scan_upto = E->d_scan_lo = E->d_scan_hi = nullptr;
    E->scan_cap = 0;
    FVM_HIP(ctx, hipMalloc((void **)&E->d_scan_upto, sizeof(int32_t) * (ntiles + 1)));
    FVM_HIP(ctx, hipMalloc((void **)&E->d_scan_lo, sizeof(int32_t) * (ntiles + 1)));
    FVM_HIP(ctx, hipMalloc((void **)&E->d_scan_hi, sizeof(int32_t) * (ntiles + 1)));
    E->scan_cap = (size_t)ntiles;
  }
  // chunk boundaries: never across a phase boundary
  std::vector<int64_t> chunk;
  for (int ph = 0; ph < PS_N; ph++) {
    const int64_t cs = ph == PS_SCAN ? 1 : (ph == PS_ORDER ? std::max(1, PS_CHUNK / 8) : PS_CHUNK);
    for (int64_t it = A.off[ph]; it < A.off[ph + 1]; it += cs) chunk.push_back(it);
  }
  chunk.push_back(A.off[PS_N]);
  A.nchunk = (int64_t)chunk.size() - 1;
  if ((size_t)chunk.size() > E->chunk_cap) {
    FVM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    free_dev(E->d_chunk);
    E->d_chunk = nullptr;
    E->chunk_cap = 0;
    FVM_HIP(ctx, hipMalloc((void **)&E->d_chunk, sizeof(int64_t) * chunk.size()));
    E->chunk_cap = chunk.size();
  }
  FVM_HIP(ctx, hipMemcpyAsync(E->d_chunk, chunk.data(), sizeof(int64_t) * chunk.size(), hipMemcpyHostToDevice,
                              ctx->stream));
  A.chunk = E->d_chunk;
  if (!E->d_done) {
    FVM_HIP(ctx, hipMalloc((void **)&E->d_pa, sizeof(PersistArgs)));
    FVM_HIP(ctx, hipMalloc((void **)&E->d_done, sizeof(int32_t) * PS_N * n));
    FVM_HIP(ctx, hipMalloc((void **)&E->d_need, sizeof(int32_t) * PS_N * n));
    FVM_HIP(ctx, hipMalloc((void **)&E->d_ticket, sizeof(unsigned long long) * 2));
  }
  FVM_HIP(ctx, hipMemcpyAsync(E->d_need, need.data(), sizeof(int32_t) * PS_N * n, hipMemcpyHostToDevice, ctx->stream));
  FVM_HIP(ctx, hipMemcpyAsync(E->d_scan_upto, upto.data(), sizeof(int32_t) * ntiles, hipMemcpyHostToDevice, ctx->stream));
  FVM_HIP(ctx, hipMemcpyAsync(E->d_scan_lo, lo.data(), sizeof(int32_t) * ntiles, hipMemcpyHostToDevice, ctx->stream));
  FVM_HIP(ctx, hipMemcpyAsync(E->d_scan_hi, hi.data(), sizeof(int32_t) * ntiles, hipMemcpyHostToDevice, ctx->stream));
  FVM_HIP(ctx, hipMemsetAsync(E->d_done, 0, sizeof(int32_t) * PS_N * n, ctx->stream));
  FVM_HIP(ctx, hipMemsetAsync(E->d_ticket, 0, sizeof(unsigned long long) * 2, ctx->stream));
  A.insts = E->dinst;
  A.n = n;
  A.ng = E->ng;
  A.nrow = E->ng;
  A.ncol = E->ncol;
  A.mblk = blk_row(E, 0), A.mbi = bi_row(E, 0);
  A.nblk = blk_row(E, 1), A.nbi = bi_row(E, 1);
  A.sblk = blk_row(E, 2 + E->ng), A.sbi = bi_row(E, 2 + E->ng);
  A.ublk = blk_row(E, u), A.ubi = bi_row(E, u);
  A.cell_off = E->dcell_off;
  A.tile_off = E->dcell_off + n + 1;
  A.tile_inst = E->dtile_inst;
  A.scan_upto = E->d_scan_upto, A.scan_lo = E->d_scan_lo, A.scan_hi = E->d_scan_hi;
  A.done = E->d_done, A.need = E->d_need, A.ticket = E->d_ticket;
  A.key = E->key, A.slot = E->slot, A.perm = E->perm, A.count = E->count, A.start = E->start;
  A.counters = E->counters;
  A.scan_fail = (int32_t *)ctx->red.ptr + FVM_FLAG_SCAN_FAIL;
  A.scan_state = (unsigned long long *)E->sums;
  A.nbucket = E->nbucket, A.nstate = nstate, A.nmark = E->nmark;
  A.xs = E->xs, A.ys = E->ys, A.pcols = E->pcols, A.part = E->part, A.dTsub = E->dTsub, A.deltaT = E->deltaT;
  A.mt = E->pcols + (int64_t)E->mt_role * E->nmark;
  A.ms = E->ms;
  A.colsrc = E->dcolsrc;
  A.groups_sorted = E->dgroups_sorted, A.groups = E->dgroups;
  A.partoff = E->dpartoff;
  A.nodes = E->dnodes;
  for (int g = 0; g < E->ng; g++) A.kg.gi[g] = g, A.kg.sx[g] = E->gstag_x[g], A.kg.sy[g] = E->gstag_y[g];
  // resident workgroups: what the kernel's registers and LDS allow on every CU
  int per_cu = 0, cus = 0;
#define PS_OCC(NF_, G_) \
  FVM_HIP(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kb_persist<NF_, G_>, 256, E->lds_locate))
  PS_DISPATCH(E->persist_nf, E->tune.lanes, PS_OCC);
#undef PS_OCC
  FVM_HIP(ctx, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
  E->persist_grid = (int)std::max<int64_t>(1, std::min<int64_t>((int64_t)std::max(per_cu, 1) * cus, A.nchunk));
  FVM_HIP(ctx, hipMemcpyAsync(E->d_pa, &A, sizeof(PersistArgs), hipMemcpyHostToDevice, ctx->stream));
  FVM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return FVM_OK;
}

static int launch_persist(fvm_ctx *ctx, fvm_ensemble *E) {
  FVM_PHASE(ctx, FVM_PHASE_PERSISTENT);
#define PS_LAUNCH(NF_, G_) \
  kb_persist<NF_, G_><<<(unsigned)E->persist_grid, 256, E->lds_locate, ctx->stream>>>(E->d_pa)
  PS_DISPATCH(E->persist_nf, E->tune.lanes, PS_LAUNCH);
#undef PS_LAUNCH
  FVM_LAUNCH_CHECK(ctx);
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
    case 4:
      kb_partials<METHOD, NF, 4><<<grid, 256, 0, ctx->stream>>>(E->dinst, n, ub, ubi, E->dgroups_sorted, E->dpartoff,
                                                                E->start, E->xs, E->ys, E->ms, E->part, kg, nrow);
      break;
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
  // the dTsub pass writes one value per marker at its store index: with fewer
  // than 8 lanes per cell those writes split into more partial lines, so it
  // takes 8 lanes when the projections take 4 (as the single-instance pass)
  switch (E->tune.lanes) {
    case 4:
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
  if (E->tune.persistent) return launch_persist(ctx, E);
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
    kb_order_gather<<<fvm_blocks(E->ntile * 32 * 64, 256), 256, 0, ctx->stream>>>(
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

extern "C" int fvm_ensemble_destroy(fvm_ensemble *E) {
  if (!E) return FVM_OK;
  if (E->exec) (void)hipGraphExecDestroy(E->exec);
  if (E->graph) (void)hipGraphDestroy(E->graph);
  void *bufs[] = {E->dinst, E->dgroups, E->dnodes, E->dblk,    E->dpartoff, E->key,       E->slot,
                  E->perm,  E->count,   E->start,  E->sums,    E->counters, E->part,      E->dTsub,
                  E->deltaT, E->xs,     E->ys,     E->pcols,   E->ms,       (void *)E->dcolsrc, E->dcell_off,
                  E->dgroups_sorted, E->dbi, E->dtile_inst, E->dbi_off, E->d_done, E->d_need, E->d_ticket,
                  E->d_scan_upto, E->d_scan_lo, E->d_scan_hi, E->d_pa, E->d_chunk};
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
  // the persistent step runs one partials kind: every group ARITHMETIC with
  // the same stored field count (at most 5)
  E->persist_nf = E->ng <= KB_MAX_MULTI ? E->gsnf[0] : 0;
  for (int gi = 0; gi < E->ng; gi++)
    if (E->gmethod[gi] != PM_ARITH || E->gsnf[gi] != E->gsnf[0] || E->gsnf[gi] > 5) E->persist_nf = 0;

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
  FVM_HIP(ctx, hipMemsetAsync(E->counters, 0, sizeof(int32_t) * 4, ctx->stream));  // [2]: sticky stall flag
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
  E->hcell = hcell;
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
  if ((t->block != 128 && t->block != 256 && t->block != 512) ||
      (t->lanes != 4 && t->lanes != 8 && t->lanes != 16 && t->lanes != 32) || (t->persistent != 0 && t->persistent != 1))
    return fvm_set_err(ctx, FVM_ERR_ARG, "tuning: block in {128,256,512}, lanes in {4,8,16,32}, persistent 0 or 1");
  if (t->persistent && t->lanes == 4)
    return fvm_set_err(ctx, FVM_ERR_ARG, "tuning: the persistent step takes lanes 8, 16 or 32");
  if (t->persistent && (t->block != 256 || E->persist_nf == 0))
    return fvm_set_err(ctx, FVM_ERR_ARG,
                       "tuning: the persistent step needs block 256 and one ARITHMETIC projection kind of <= 5 fields");
  if (E->tune.block == t->block && E->tune.lanes == t->lanes && E->tune.persistent == t->persistent) return FVM_OK;
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
  const int blocks[4] = {128, 256, 512, 256}, lanes[4] = {8, 16, 32, 4};
  fvm_tuning bt = E->tune;
  float bms = 1e30f;
  // per-phase launches with every block size and lane count, then the
  // persistent step (if the projection list allows it; not with 4 lanes)
  for (int bi = 0; bi < (E->persist_nf ? 4 : 3); bi++)
    for (int li = 0; li < (bi == 3 ? 3 : 4); li++) {
      fvm_tuning t = {blocks[bi], lanes[li], bi == 3 ? 1 : 0};
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
  int32_t h[3] = {0, 0, 0};
  FVM_HIP(ctx, hipMemcpyAsync(h, E->counters, 3 * sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
  FVM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (nOut) *nOut = h[0];
  if (nBad) *nBad = h[1];
  if (h[2]) return fvm_set_err(ctx, FVM_ERR_STALLED, "a persistent ensemble step gave up waiting; its results are incomplete");
  return FVM_OK;
}
