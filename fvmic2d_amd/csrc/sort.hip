// Marker cell location (findMarkerCells) and the deterministic bucket order
// used by every projection.
//
// Bucket key of an active marker: ((cellY*ncx + cellX) << 2) | bx | (by << 1),
// ncx = NX - 1 basic cells along x (NX on a periodic grid),
// with bx = X >= xc[cellX+1], by = Y >= yc[cellY+1] -- the quadrant tests of
// findCellStag2D (markerProjection.c:258-268).  For every staggered node family
// all markers of one bucket fall into the same family cell, so a family cell is
// the union of at most four buckets.  Buckets are filled with integer atomics
// and then each bucket is put in ascending marker-index order, so the final
// order is a pure function of the keys (identical to a stable sort by key).
#include "fvm_internal.h"
#include "scan_common.h"

// ---- findMarkerCell (markers.c:41-83) ---------------------------------------
// The cell index along each axis is the bisection's (markers.c:48-59), found
// from a guess (guess_grid) -- the same index for any non-decreasing grid.

// Cell location, fused with what the step does next: the bucket key and its
// count (when `count` is non-null) and the density loop of main_subduction.c:
// 255-264 (when `dens` is non-null).  Counts use one atomic per distinct key
// per wavefront; the slot they return is only a provisional position inside
// the bucket -- k_bucket_order fixes the final order.
struct DensityTab {
  double rho[FVM_MAXMAT], alpha[FVM_MAXMAT];
};

#define LOC_TILE_BYTES 4096
template <bool LDS>
__global__ void __launch_bounds__(256) k_locate(DGrid g, int32_t n, const double *__restrict__ X,
                                                const double *__restrict__ Y, int32_t *__restrict__ cellX,
                                                int32_t *__restrict__ cellY, int32_t *__restrict__ cpu, int32_t rank,
                                                int32_t *__restrict__ counters, int32_t *__restrict__ key,
                                                int32_t *__restrict__ slot, int32_t *__restrict__ count,
                                                DensityTab dens, int do_dens, const int32_t *__restrict__ Mat,
                                                const double *__restrict__ T, double *__restrict__ rho,
                                                double *__restrict__ rhodot, int tiles) {
  extern __shared__ double sgrid[];
  const double *gx = g.x, *gy = g.y, *gxc = g.xc, *gyc = g.yc;
  if (LDS) {
    // x[0..nxn-1], y[0..NY-1], xc[0..NX], yc[0..NY]; nxn = NX + per nodes
    // bound the cells along x (x[NX] = LX closes a periodic grid's last cell)
    const int nxn = g.NX + g.per;
    for (int i = threadIdx.x; i < nxn; i += blockDim.x) sgrid[i] = g.x[i];
    for (int i = threadIdx.x; i < g.NY; i += blockDim.x) sgrid[nxn + i] = g.y[i];
    for (int i = threadIdx.x; i <= g.NX; i += blockDim.x) sgrid[nxn + g.NY + i] = g.xc[i];
    for (int i = threadIdx.x; i <= g.NY; i += blockDim.x) sgrid[nxn + g.NX + g.NY + 1 + i] = g.yc[i];
    __syncthreads();
    gx = sgrid;
    gy = sgrid + nxn;
    gxc = sgrid + nxn + g.NY;
    gyc = sgrid + nxn + g.NX + g.NY + 1;
  }
  // findMarkerCell bisects x over [0, NX - 1], or [0, NX] on a periodic grid
  // (markers.c:49), and accepts cells up to NX - 2, or NX - 1 (markers.c:73)
  const int ncx = g.ncx();
  const double ginvx = (double)ncx / (gx[ncx] - gx[0]);
  const double ginvy = (double)(g.NY - 1) / (gy[g.NY - 1] - gy[0]);
  // each block locates `tiles` consecutive tiles of blockDim.x markers, so that
  // the grid staged in LDS serves more than one tile
  int nout = 0, nbad = 0;
  for (int t = 0; t < tiles; t++) {
    const int64_t m = ((int64_t)blockIdx.x * tiles + t) * blockDim.x + threadIdx.x;
    int out = 0, bad = 0, k = -1;
    if (m < n) {
      const double xm = X[m], ym = Y[m];
      int cx = -1, cy = -1;
      if (xm >= 0.0 && ym >= 0.0 && xm <= g.LX && ym <= g.LY) {
        cx = guess_grid(gx, ncx + 1, xm, ginvx);
        cy = guess_grid(gy, g.NY, ym, ginvy);
        if (cx < 0 || cx > ncx - 1 || cy < 0 || cy > g.NY - 2) {
          bad = 1;
        } else {
          const int bx = xm >= gxc[cx + 1] ? 1 : 0;
          const int by = ym >= gyc[cy + 1] ? 1 : 0;
          k = ((cy * ncx + cx) << 2) | bx | (by << 1);
        }
      } else {
        out = 1;
      }
      cellX[m] = cx;
      cellY[m] = cy;
      cpu[m] = rank;
      if (do_dens && cx != -1) {
        const int mm = Mat[m] & (FVM_MAXMAT - 1);
        rho[m] = -dens.rho[mm] * (dens.alpha[mm] * T[m]);
        rhodot[m] = 0.0;
      }
      if (count) key[m] = k;
    }
    if (count) {
      const int sl = wave_count_slot(k, count);
      if (k >= 0) slot[m] = sl;
    }
    nout += __popcll(__ballot(out));
    nbad += __popcll(__ballot(bad));
  }
  // counts of outside markers and bad cells (integer: order-free), one atomic
  // per wavefront that has any
  if ((threadIdx.x & 63) == 0) {
    if (nout) atomicAdd(&counters[0], nout);
    if (nbad) atomicAdd(&counters[1], nbad);
  }
}

// ---- bucket order ----------------------------------------------------------------
// Bucket counts of markers whose cells are already known (a projection after a
// caller-side findMarkerCells): same wavefront-aggregated atomics as k_locate.
__global__ void __launch_bounds__(256) k_keys(DGrid g, int32_t n, const double *__restrict__ X,
                                              const double *__restrict__ Y, const int32_t *__restrict__ cellX,
                                              const int32_t *__restrict__ cellY, int32_t *__restrict__ key,
                                              int32_t *__restrict__ slot, int32_t *__restrict__ count) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int k = -1;
  if (m < n) {
    const int cx = cellX[m], cy = cellY[m];
    if (cell_ok(cx, cy, g.ncx(), g.NY)) {  // inactive (or invalid) markers take no part
      const int bx = X[m] >= g.xc[cx + 1] ? 1 : 0;
      const int by = Y[m] >= g.yc[cy + 1] ? 1 : 0;
      k = ((cy * g.ncx() + cx) << 2) | bx | (by << 1);
    }
    key[m] = k;
  }
  const int sl = wave_count_slot(k, count);
  if (k >= 0) slot[m] = sl;
}

__global__ void __launch_bounds__(SCAN_BS) k_scan_onepass(const int32_t *__restrict__ in, int32_t *__restrict__ out,
                                                          unsigned long long *__restrict__ state, int64_t n,
                                                          int32_t *__restrict__ fail) {
  __shared__ int32_t s_tile;
  if (threadIdx.x == 0) s_tile = (int32_t)atomicAdd(&state[0], 1ull);
  __syncthreads();
  scan_tile(s_tile, in, out, state, n, fail);
}

// counts (n entries, optional), the two location counters and the scan state
__global__ void __launch_bounds__(256) k_zero_counts(int32_t *__restrict__ count, int64_t n,
                                                     int32_t *__restrict__ counters,
                                                     unsigned long long *__restrict__ state, int64_t nstate) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (count && i < n) count[i] = 0;
  if (state && i < nstate) state[i] = 0ull;
  if (counters && i < 2) counters[i] = 0;
}

__global__ void __launch_bounds__(256) k_scatter(int32_t n, const int32_t *__restrict__ key,
                                                 const int32_t *__restrict__ slot, const int32_t *__restrict__ start,
                                                 int32_t *__restrict__ perm) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n) return;
  const int k = key[m];
  if (k < 0) return;
  const int32_t p = start[k] + slot[m];
  if ((uint32_t)p < (uint32_t)n) perm[p] = (int32_t)m;  // always true for a valid scan
}

// Put every bucket in ascending marker order (bucket_order_pair: two basic
// cells per wavefront).  With GATHER the payload columns of each entry are
// copied to its final position as well, and the cells are visited in 8 x 8
// tiles (y fastest inside a tile; a wavefront takes two cells one above the
// other), tiles spread over the XCDs in contiguous runs, so that the marker
// rows a tile's wavefronts read share L2 lines.
template <bool GATHER>
__global__ void __launch_bounds__(256) k_bucket_order(int32_t ncell, int32_t ncx, int32_t ncy,
                                                      const int32_t *__restrict__ start, int32_t *__restrict__ perm,
                                                      SortGather sg) {
  int64_t c0, c1;
  if (GATHER) {
    const int64_t w = xcd_tile(blockIdx.x, gridDim.x) * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const uint32_t ntx = (uint32_t)(ncx + 7) >> 3;
    const uint32_t tile = (uint32_t)(w >> 5);
    const int l = (int)(w & 31) * 2;
    const int cx = (int)(tile % ntx) * 8 + (l >> 3), cy = (int)(tile / ntx) * 8 + (l & 7);
    const bool in = cx < ncx && tile / ntx < (uint32_t)((ncy + 7) >> 3);
    c0 = (in && cy < ncy) ? (int64_t)cy * ncx + cx : -1;
    c1 = (in && cy + 1 < ncy) ? (int64_t)(cy + 1) * ncx + cx : -1;
  } else {
    const int64_t w = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    c0 = 2 * w < ncell ? 2 * w : -1;
    c1 = 2 * w + 1 < ncell ? 2 * w + 1 : -1;
  }
  if (c0 < 0 && c1 < 0) return;  // whole wavefront exits together
  bucket_order_pair(c0, c1, start, perm, [&](int p, int v) {
    if (GATHER) gather_entry(sg.pc, sg.with_xy, sg.X, sg.Y, sg.Mat, sg.xs, sg.ys, sg.ms, p, v);
  });
}

static int sort_buffers(fvm_ctx *ctx, const fvm_grid *g, int32_t n) {
  const int64_t nk = 4LL * fvm_ncx(g) * (g->NY - 1);
  const int64_t ntiles = (nk + SCAN_TILE - 1) / SCAN_TILE;
  FVM_TRY(fvm_ws(ctx, ctx->key, (size_t)(n > 0 ? n : 1) * 4));
  FVM_TRY(fvm_ws(ctx, ctx->slot, (size_t)(n > 0 ? n : 1) * 4));
  FVM_TRY(fvm_ws(ctx, ctx->perm, (size_t)(n > 0 ? n : 1) * 4));
  FVM_TRY(fvm_ws(ctx, ctx->count, (size_t)(nk + 1) * 4));
  FVM_TRY(fvm_ws(ctx, ctx->start, (size_t)(nk + 1) * 4));
  FVM_TRY(fvm_ws(ctx, ctx->blocksums, (size_t)(ntiles + 2) * 8));  // one-pass scan state
  return FVM_OK;
}

// Zero the bucket counts, the one-pass scan state and the location counters
// (one launch); without a grid only the counters.
static int zero_counts(fvm_ctx *ctx, const fvm_grid *g) {
  int32_t *counters = (int32_t *)ctx->red.ptr;
  FVM_PHASE(ctx, FVM_PHASE_SORT);
  if (!g) {
    k_zero_counts<<<1, 64, 0, ctx->stream>>>(nullptr, 0, counters, nullptr, 0);
  } else {
    const int64_t nk = 4LL * fvm_ncx(g) * (g->NY - 1);
    const int64_t ntiles = (nk + SCAN_TILE - 1) / SCAN_TILE;
    k_zero_counts<<<fvm_blocks(nk + 1, 256), 256, 0, ctx->stream>>>(
        (int32_t *)ctx->count.ptr, nk + 1, counters, (unsigned long long *)ctx->blocksums.ptr, ntiles + 2);
  }
  FVM_LAUNCH_CHECK(ctx);
  return FVM_OK;
}

// scan + scatter + in-bucket order, after the counts and provisional slots exist
static int finish_sort(fvm_ctx *ctx, const fvm_grid *g, const fvm_markers *mk, const SortGather *sg) {
  FVM_PHASE(ctx, FVM_PHASE_SORT);
  const int32_t n = mk->nMark;
  const int64_t ncell = (int64_t)fvm_ncx(g) * (g->NY - 1);
  const int64_t nk = 4 * ncell;
  const int64_t ntiles = (nk + SCAN_TILE - 1) / SCAN_TILE;
  int32_t *key = (int32_t *)ctx->key.ptr, *slot = (int32_t *)ctx->slot.ptr, *perm = (int32_t *)ctx->perm.ptr;
  int32_t *count = (int32_t *)ctx->count.ptr, *start = (int32_t *)ctx->start.ptr;
  k_scan_onepass<<<(unsigned)ntiles, SCAN_BS, 0, ctx->stream>>>(
      count, start, (unsigned long long *)ctx->blocksums.ptr, nk, (int32_t *)ctx->red.ptr + FVM_FLAG_SCAN_FAIL);
  FVM_LAUNCH_CHECK(ctx);
  if (n > 0) {
    k_scatter<<<fvm_blocks(n, 256), 256, 0, ctx->stream>>>(n, key, slot, start, perm);
    FVM_LAUNCH_CHECK(ctx);
  }
  if (sg && n > 0) {
    const int32_t ncx = fvm_ncx(g), ncy = g->NY - 1;
    const int64_t ntiles = (int64_t)((ncx + 7) / 8) * ((ncy + 7) / 8);
    k_bucket_order<true><<<fvm_blocks(ntiles * 32 * 64, 256), 256, 0, ctx->stream>>>((int32_t)ncell, ncx, ncy, start,
                                                                                   perm, *sg);
  } else {
    SortGather none;
    memset(&none, 0, sizeof(none));
    k_bucket_order<false><<<fvm_blocks((ncell + 1) / 2 * 64, 256), 256, 0, ctx->stream>>>((int32_t)ncell, 0, 0, start,
                                                                                          perm, none);
  }
  FVM_LAUNCH_CHECK(ctx);
  ctx->sort_epoch = ctx->epoch;
  ctx->sort_serial++;
  ctx->sort_x = g->x;
  ctx->sort_X = mk->X;
  ctx->sort_n = n;
  return FVM_OK;
}

int fvm_locate_async(fvm_ctx *ctx, const fvm_grid *g, fvm_markers *mk, int32_t rank, const Materials *dens,
                     bool build_order) {
  int32_t *counters = (int32_t *)ctx->red.ptr;
  const int32_t n = mk->nMark;
  fvm_markers_moved(ctx);
  if (build_order) FVM_TRY(sort_buffers(ctx, g, n));
  FVM_TRY(zero_counts(ctx, build_order ? g : nullptr));
  DensityTab dt;
  for (int i = 0; i < FVM_MAXMAT; i++) {
    dt.rho[i] = dens ? dens->materialRho[i] : 0.0;
    dt.alpha[i] = dens ? dens->materialAlpha[i] : 0.0;
  }
  if (n > 0) {
    FVM_PHASE(ctx, FVM_PHASE_LOCATE);
    const DGrid dg = dgrid(g);
    int32_t *key = build_order ? (int32_t *)ctx->key.ptr : nullptr;
    int32_t *slot = build_order ? (int32_t *)ctx->slot.ptr : nullptr;
    int32_t *count = build_order ? (int32_t *)ctx->count.ptr : nullptr;
    const size_t lds = (size_t)(2 * g->NX + (g->xperiodic ? 1 : 0) + 2 * g->NY + 2) * sizeof(double);
    if (lds <= 48 * 1024) {
      // one tile of 256 markers per LOC_TILE_BYTES of staged grid, at least 2
      // (at 201 x 201, 6.4 KB staged, 2 tiles locate in 37 us against 42 for 1)
      // and at most 8
      const int tiles = std::max(2, std::min(8, (int)(lds / LOC_TILE_BYTES)));
      k_locate<true><<<fvm_blocks(n, 256 * tiles), 256, lds, ctx->stream>>>(
          dg, n, mk->X, mk->Y, mk->cellX, mk->cellY, mk->cpu, rank, counters, key, slot, count, dt, dens ? 1 : 0,
          mk->Mat, mk->T, mk->rho, mk->rhodot, tiles);
    } else {
      k_locate<false><<<fvm_blocks(n, 256), 256, 0, ctx->stream>>>(dg, n, mk->X, mk->Y, mk->cellX, mk->cellY,
                                                                    mk->cpu, rank, counters, key, slot, count, dt,
                                                                    dens ? 1 : 0, mk->Mat, mk->T, mk->rho,
                                                                    mk->rhodot, 1);
    }
    FVM_LAUNCH_CHECK(ctx);
  }
  if (build_order) {  // finished by the first projection (fvm_sort_markers)
    ctx->pend = true;
    ctx->pend_x = g->x;
    ctx->pend_X = mk->X;
    ctx->pend_n = n;
    ctx->pend_epoch = ctx->epoch;
  }
  return FVM_OK;
}

extern "C" int fvm_find_marker_cells(fvm_ctx *ctx, const fvm_grid *g, fvm_markers *mk, int32_t rank,
                                     int32_t *nOut) {
  if (!ctx || !mk) return FVM_ERR_ARG;
  FVM_TRY(fvm_grid_check(ctx, g));
  FVM_TRY(fvm_locate_async(ctx, g, mk, rank, nullptr, false));
  int32_t *counters = (int32_t *)ctx->red.ptr;
  int32_t *h = (int32_t *)ctx->pinned.ptr;
  FVM_HIP(ctx, hipMemcpyAsync(h, counters, 2 * sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
  FVM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (nOut) *nOut = h[0];
  if (h[1]) return fvm_set_err(ctx, FVM_ERR_CELL_OUT_OF_BOUNDS, "%d markers located outside the cell range", h[1]);
  return FVM_OK;
}

bool fvm_sort_valid(const fvm_ctx *ctx, const fvm_grid *g, const fvm_markers *mk) {
  return ctx->sort_epoch == ctx->epoch && ctx->sort_x == (const void *)g->x && ctx->sort_X == (const void *)mk->X &&
         ctx->sort_n == mk->nMark;
}

int fvm_sort_markers(fvm_ctx *ctx, const fvm_grid *g, const fvm_markers *mk, const SortGather *sg,
                     bool *gathered) {
  if (gathered) *gathered = false;
  if (fvm_sort_valid(ctx, g, mk)) return FVM_OK;
  if (ctx->pend && ctx->pend_epoch == ctx->epoch && ctx->pend_x == (const void *)g->x &&
      ctx->pend_X == (const void *)mk->X && ctx->pend_n == mk->nMark) {
    ctx->pend = false;
    FVM_TRY(finish_sort(ctx, g, mk, sg));
    if (gathered) *gathered = sg != nullptr && mk->nMark > 0;
    return FVM_OK;
  }
  ctx->pend = false;
  const int32_t n = mk->nMark;
  FVM_TRY(sort_buffers(ctx, g, n));
  FVM_TRY(zero_counts(ctx, g));
  if (n > 0) {
    FVM_PHASE(ctx, FVM_PHASE_SORT);
    k_keys<<<fvm_blocks(n, 256), 256, 0, ctx->stream>>>(dgrid(g), n, mk->X, mk->Y, mk->cellX, mk->cellY,
                                                        (int32_t *)ctx->key.ptr, (int32_t *)ctx->slot.ptr,
                                                        (int32_t *)ctx->count.ptr);
    FVM_LAUNCH_CHECK(ctx);
  }
  FVM_TRY(finish_sort(ctx, g, mk, sg));
  if (gathered) *gathered = sg != nullptr && n > 0;
  return FVM_OK;
}

// ---- store reorder (fvm_markers_reorder) ---------------------------------------------
// Markers outside the bucket order (key < 0: outside the domain or not located)
// are flagged, and an exclusive scan of the flags (k_scan_onepass) gives their
// rank among themselves, so they follow the ordered markers in index order.
__global__ void __launch_bounds__(256) k_flag_unordered(int32_t n, const int32_t *__restrict__ key,
                                                        int32_t *__restrict__ flag) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m < n) flag[m] = key[m] < 0 ? 1 : 0;
}

// order[p] = index of the marker that moves to slot p: the bucket order for
// p < nact = *total (the scan's total of the bucket counts), then the
// unordered markers by index.
__global__ void __launch_bounds__(256) k_reorder_index(int32_t n, const int32_t *__restrict__ key,
                                                       const int32_t *__restrict__ prefix,
                                                       const int32_t *__restrict__ total,
                                                       const int32_t *__restrict__ perm, int32_t *__restrict__ order) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n) return;
  const int32_t nact = *total;
  if (m < nact) order[m] = perm[m];
  if (key[m] < 0) {
    const int32_t p = nact + prefix[m];
    if ((uint32_t)p < (uint32_t)n) order[p] = (int32_t)m;  // always true for a valid scan
  }
}

#define REORDER_COLS 8
template <class T>
struct PermTab {
  int nc;
  const T *src[REORDER_COLS];
  T *dst[REORDER_COLS];
};
// dst[k][p] = src[k][order[p]]: the staging copy of up to REORDER_COLS columns
template <class T>
__global__ void __launch_bounds__(256) k_permute_cols(int32_t n, const int32_t *__restrict__ order, PermTab<T> t) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int v = order[p];
#pragma unroll
  for (int k = 0; k < REORDER_COLS; k++)
    if (k < t.nc) t.dst[k][p] = t.src[k][v];
}

static size_t align256(size_t b) { return (b + 255) / 256 * 256; }

template <class T>
static int permute_columns(fvm_ctx *ctx, int32_t n, const int32_t *order, T **cols, int ncols, char *stage) {
  for (int c0 = 0; c0 < ncols; c0 += REORDER_COLS) {
    PermTab<T> t;
    t.nc = std::min(REORDER_COLS, ncols - c0);
    for (int k = 0; k < t.nc; k++) {
      t.src[k] = cols[c0 + k];
      t.dst[k] = (T *)(stage + align256((size_t)n * sizeof(T)) * k);
    }
    k_permute_cols<T><<<fvm_blocks(n, 256), 256, 0, ctx->stream>>>(n, order, t);
    FVM_LAUNCH_CHECK(ctx);
    for (int k = 0; k < t.nc; k++)
      FVM_HIP(ctx, hipMemcpyAsync(cols[c0 + k], t.dst[k], (size_t)n * sizeof(T), hipMemcpyDeviceToDevice,
                                  ctx->stream));
  }
  return FVM_OK;
}

extern "C" int fvm_markers_reorder(fvm_ctx *ctx, const fvm_grid *g, fvm_markers *mk, int32_t *order) {
  if (!ctx || !mk) return FVM_ERR_ARG;
  FVM_TRY(fvm_grid_check(ctx, g));
  const int32_t n = mk->nMark;
  if (n <= 0) return FVM_OK;
  FVM_TRY(fvm_sort_markers(ctx, g, mk));
  FVM_PHASE(ctx, FVM_PHASE_SORT);
  const int64_t nk = 4LL * fvm_ncx(g) * (g->NY - 1);
  const int64_t tiles = (n + SCAN_TILE - 1) / SCAN_TILE;
  const size_t b_flag = align256((size_t)n * 4), b_prefix = align256((size_t)(n + 1) * 4);
  const size_t b_state = align256((size_t)(tiles + 2) * 8), b_order = align256((size_t)n * 4);
  const size_t b_stage = align256((size_t)n * 8) * REORDER_COLS;
  FVM_TRY(fvm_ws(ctx, ctx->reord, b_flag + b_prefix + b_state + b_order + b_stage));
  char *w = (char *)ctx->reord.ptr;
  int32_t *flag = (int32_t *)w, *prefix = (int32_t *)(w + b_flag);
  unsigned long long *state = (unsigned long long *)(w + b_flag + b_prefix);
  int32_t *ord = order ? order : (int32_t *)(w + b_flag + b_prefix + b_state);
  char *stage = w + b_flag + b_prefix + b_state + b_order;
  const int32_t *key = (const int32_t *)ctx->key.ptr, *perm = (const int32_t *)ctx->perm.ptr;
  const int32_t *total = (const int32_t *)ctx->start.ptr + nk;
  k_zero_counts<<<fvm_blocks(tiles + 2, 256), 256, 0, ctx->stream>>>(nullptr, 0, nullptr, state, tiles + 2);
  FVM_LAUNCH_CHECK(ctx);
  k_flag_unordered<<<fvm_blocks(n, 256), 256, 0, ctx->stream>>>(n, key, flag);
  FVM_LAUNCH_CHECK(ctx);
  k_scan_onepass<<<(unsigned)tiles, SCAN_BS, 0, ctx->stream>>>(flag, prefix, state, n,
                                                               (int32_t *)ctx->red.ptr + FVM_FLAG_SCAN_FAIL);
  FVM_LAUNCH_CHECK(ctx);
  k_reorder_index<<<fvm_blocks(n, 256), 256, 0, ctx->stream>>>(n, key, prefix, total, perm, ord);
  FVM_LAUNCH_CHECK(ctx);
  double *dc[FVM_NUM_DCOLS];
  int32_t *ic[FVM_NUM_ICOLS];
  fvm_marker_columns(mk, dc, ic);
  FVM_TRY(permute_columns<double>(ctx, n, ord, dc, FVM_NUM_DCOLS, stage));
  FVM_TRY(permute_columns<int32_t>(ctx, n, ord, ic, FVM_NUM_ICOLS, stage));
  fvm_markers_moved(ctx);
  return FVM_OK;
}

// ---- helpers for the batched ensemble (ensemble.hip) ---------------------------------
int64_t fvm_scan_tiles(int64_t nk) { return (nk + SCAN_TILE - 1) / SCAN_TILE; }

int fvm_launch_zero_sort(fvm_ctx *ctx, int32_t *count, int64_t ncount, int32_t *counters, unsigned long long *state,
                         int64_t nstate) {
  FVM_PHASE(ctx, FVM_PHASE_SORT);
  const int64_t nmax = ncount > nstate ? ncount : nstate;
  k_zero_counts<<<fvm_blocks(nmax > 2 ? nmax : 2, 256), 256, 0, ctx->stream>>>(count, ncount, counters, state,
                                                                              nstate);
  FVM_LAUNCH_CHECK(ctx);
  return FVM_OK;
}

int fvm_launch_scan_onepass(fvm_ctx *ctx, const int32_t *count, int32_t *start, unsigned long long *state,
                            int64_t nk) {
  FVM_PHASE(ctx, FVM_PHASE_SORT);
  k_scan_onepass<<<(unsigned)fvm_scan_tiles(nk), SCAN_BS, 0, ctx->stream>>>(count, start, state, nk,
                                                                          (int32_t *)ctx->red.ptr + FVM_FLAG_SCAN_FAIL);
  FVM_LAUNCH_CHECK(ctx);
  return FVM_OK;
}

