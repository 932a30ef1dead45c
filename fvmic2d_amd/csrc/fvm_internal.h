// Internal declarations shared by the HIP translation units of libfvmic2d_hip.
// All device code is compiled with -ffp-contract=off: the position and cell
// chain of the marker path must round exactly like the reference's gcc build
// (no fused multiply-add), so every expression below keeps the reference's
// operation order.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "fvm_abi.h"

struct fvm_buf {
  void *ptr = nullptr;
  size_t bytes = 0;
};

struct fvm_ctx {
  hipStream_t stream = nullptr;
  bool own_stream = false;
  int device = 0;
  char err[512] = {0};
  // workspace buffers, grown on demand
  fvm_buf key, slot, perm, count, start, fill, blocksums, partials, scratch, staging, red;
  fvm_buf xfer;  // packed nodal arrays of fvm_nodal_upload / fvm_nodal_download
  fvm_buf popc, popc2;  // population control: per-quarter plan and out list; added markers
  fvm_buf payload;     // bucket-ordered copies of the marker columns a projection reads
  fvm_buf payload_xy;  // bucket-ordered X, Y, Mat: valid while xy_serial == sort_serial
  fvm_buf reord;       // fvm_markers_reorder: flags, prefix, scan state, order, column staging
  uint64_t sort_serial = 0, xy_serial = 0;
  // columns held in `payload` (slot k = payload + k * stride doubles) for sort
  // payload_serial; a fused subgrid pass may take its T column from there only
  // while reuse_payload is set by a caller that knows the column is unchanged
  const double *payload_src[24] = {nullptr};
  int payload_nc = 0;
  size_t payload_stride = 0;
  uint64_t payload_serial = 0;
  bool reuse_payload = false;
  int lanes = 8;  // lanes per lattice cell in the projection partials (4, 8, 16 or 32)
  int jaumann = 0;  // stress rotation of the whole-step entry points
  fvm_buf pinned;  // small pinned host buffer for scalar results
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // optional per-phase timing (fvm_profile_*): event pairs around launches
  int profiling = 0;
  int nprof = 0;
  hipEvent_t prof_ev[2 * 1024];
  int prof_phase[1024];
  // sort cache: the bucket order is valid for (grid, markers, epoch)
  const void *sort_x = nullptr;
  const void *sort_X = nullptr;
  int32_t sort_n = -1;
  uint64_t sort_epoch = 0, epoch = 1;
  int32_t sort_nactive = 0;
  // keys, slots and bucket counts left by a locate pass for (grid, markers,
  // epoch); the first projection finishes that order (and gathers its payload)
  bool pend = false;
  const void *pend_x = nullptr;
  const void *pend_X = nullptr;
  int32_t pend_n = -1;
  uint64_t pend_epoch = 0;
};

int fvm_set_err(fvm_ctx *ctx, int code, const char *fmt, ...);
int fvm_ws(fvm_ctx *ctx, fvm_buf &b, size_t bytes);  // ensure workspace capacity

#define FVM_HIP(ctx, call)                                                               \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess)                                                                \
      return fvm_set_err((ctx), FVM_ERR_HIP, "%s:%d %s: %s", __FILE__, __LINE__, #call,  \
                         hipGetErrorString(e_));                                         \
  } while (0)

#define FVM_TRY(expr)        \
  do {                       \
    int r_ = (expr);         \
    if (r_ != FVM_OK) return r_; \
  } while (0)

#define FVM_LAUNCH_CHECK(ctx) FVM_HIP(ctx, hipGetLastError())

// Marker-path entry points take periodic grids (x periodic, NX >= 4: the two
// ghost columns on each side then wrap to distinct nodes); the nodal solves and
// population control call fvm_grid_check_nonperiodic instead.
static inline int fvm_grid_check(fvm_ctx *ctx, const fvm_grid *g) {
  if (!g || !g->x || !g->y || !g->xc || !g->yc || g->NX < 2 || g->NY < 2)
    return fvm_set_err(ctx, FVM_ERR_ARG, "invalid grid descriptor");
  if (g->xperiodic && g->NX < 4) return fvm_set_err(ctx, FVM_ERR_PERIODIC, "a periodic grid needs NX >= 4");
  return FVM_OK;
}
static inline int fvm_grid_check_nonperiodic(fvm_ctx *ctx, const fvm_grid *g, const char *what) {
  FVM_TRY(fvm_grid_check(ctx, g));
  if (g->xperiodic) return fvm_set_err(ctx, FVM_ERR_PERIODIC, "%s: periodic grids are not supported", what);
  return FVM_OK;
}
// Basic cells along x: NX - 1, or NX on a periodic grid (x[NX] = LX closes the
// last cell; makePeriodic, gridSpacing.c:159-168).
static inline int fvm_ncx(const fvm_grid *g) { return g->NX - 1 + (g->xperiodic ? 1 : 0); }

static inline unsigned fvm_blocks(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// ---- XCD-aware block order ----------------------------------------------------
// Workgroups are dispatched round-robin over the 8 XCDs of an MI355X, each with
// its own L2.  Kernels whose neighbouring blocks share cache lines (projection
// passes: adjacent lattice cells read the same marker lines) renumber their
// blocks so that every XCD works on one contiguous range: block b runs on XCD
// b % 8 and becomes tile (b % 8) * q + min(b % 8, r) + b / 8, q = nb / 8,
// r = nb % 8 -- a bijection on [0, nb).
#define FVM_NUM_XCD 8

// int32 slot in ctx->red: sticky diagnostic flag set when the one-pass bucket
// scan stopped waiting on a predecessor tile and summed the counts itself
#define FVM_FLAG_SCAN_FAIL 2
__device__ __forceinline__ int64_t xcd_tile(int64_t b, int64_t nb) {
  const int64_t q = nb / FVM_NUM_XCD, r = nb % FVM_NUM_XCD;
  const int64_t x = b % FVM_NUM_XCD, k = b / FVM_NUM_XCD;
  return x * q + (x < r ? x : r) + k;
}

// A device pointer read from memory is generic to the compiler, and a load
// through it is a flat load, which also counts on the scalar/LDS counter: a
// wait for any scalar load then waits for it too.  Marker columns (device
// global memory) read through a pointer held in a struct are read through this
// cast, as global loads.
typedef __attribute__((address_space(1))) const double gdouble;
__device__ __forceinline__ const gdouble *global_ptr(const double *p) { return (const gdouble *)p; }

// Wavefront-uniform lane broadcast (v_readlane): cheaper than ds_bpermute.
__device__ __forceinline__ int read_lane(int v, int lane) { return __builtin_amdgcn_readlane(v, lane); }

// Wavefront-aggregated bucket counting: lanes with the same key k >= 0 share
// one atomicAdd on count[k]; every lane gets a provisional slot (the old count
// plus its rank among equal keys in lane order).  All leaders issue their
// atomics together, so the wavefront waits for one round trip.
__device__ __forceinline__ int wave_count_slot(int k, int32_t *__restrict__ count) {
  const int lane = threadIdx.x & 63;
  unsigned long long todo = __ballot(k >= 0);
  int leader = 0, rank = 0, npeer = 0;
  while (todo) {
    const int l = __ffsll((long long)todo) - 1;
    const int kl = read_lane(k, l);
    const unsigned long long peers = __ballot(k == kl) & todo;
    if (k == kl) {
      leader = l;
      rank = __popcll(peers & ((1ull << lane) - 1ull));
      npeer = __popcll(peers);
    }
    todo &= ~peers;
  }
  int base = 0;
  if (k >= 0 && lane == leader) base = atomicAdd(&count[k], npeer);
  base = __shfl(base, leader);
  return base + rank;
}

// ---- device-side grid view -------------------------------------------------
struct DGrid {
  int NX, NY;
  int per;  // 1: x periodic (node index ix stands for ix mod NX)
  double LX, LY;
  const double *x, *y, *xc, *yc;
  __host__ __device__ int ncx() const { return NX - 1 + per; }
};
static inline DGrid dgrid(const fvm_grid *g) {
  DGrid d;
  d.NX = g->NX;
  d.NY = g->NY;
  d.per = g->xperiodic ? 1 : 0;
  d.LX = g->LX;
  d.LY = g->LY;
  d.x = g->x;
  d.y = g->y;
  d.xc = g->xc;
  d.yc = g->yc;
  return d;
}

// A marker takes part only if its cell is inside the grid: cellX == -1 marks an
// inactive marker (fdcode.h:120); any other out-of-range index (a caller that
// skipped findMarkerCells) is skipped too, so no kernel indexes outside its arrays.
// ncx = basic cells along x (NX - 1, or NX on a periodic grid)
__device__ __forceinline__ bool cell_ok(int cx, int cy, int ncx, int NY) {
  return (unsigned)cx < (unsigned)ncx && (unsigned)cy <= (unsigned)(NY - 2);
}

// x node index of a periodic local array: ghost columns -2, -1 and NX, NX + 1
// stand for NX - 2, NX - 1 and 0, 1 (DM_BOUNDARY_PERIODIC, nodalFields.c:11-20)
__device__ __forceinline__ int wrap_x(int ix, int NX, int per) {
  return per ? (ix < 0 ? ix + NX : (ix >= NX ? ix - NX : ix)) : ix;
}

// Read of a ghosted local array: indices outside the grid read 0 (a fresh,
// zeroed DMDA local vector; DMGlobalToLocal fills only owned nodes); on a
// periodic grid the x ghosts read the wrapped column.  The load itself is
// unconditional (an index outside the grid reads element 0 and the value is
// dropped), so a kernel's gathers are issued together, without a branch each.
__device__ __forceinline__ double gat(const double *__restrict__ f, int NX, int NY, int per, int jy, int ix) {
  ix = wrap_x(ix, NX, per);
  const bool in = ix >= 0 && ix < NX && jy >= 0 && jy < NY;
  const double v = f[in ? (size_t)jy * NX + ix : 0];
  return in ? v : 0.0;
}

// The reference's bilinear gather written as four accumulations
// (updateMarkerStress.c:84-88 form): v=0; v+=(1-mdx)(1-mdy)f00; v+=(1-mdx)mdy f10; ...
__device__ __forceinline__ double bil_acc(const double *__restrict__ f, int NX, int NY, int per, int cy, int cx,
                                          double mdx, double mdy) {
  double v = 0.0;
  v += (1.0 - mdx) * (1.0 - mdy) * gat(f, NX, NY, per, cy, cx);
  v += (1.0 - mdx) * mdy * gat(f, NX, NY, per, cy + 1, cx);
  v += mdx * (1.0 - mdy) * gat(f, NX, NY, per, cy, cx + 1);
  v += mdx * mdy * gat(f, NX, NY, per, cy + 1, cx + 1);
  return v;
}

// ---- projection engine (project.hip) --------------------------------------
enum {
  PM_ARITH = 0,   // ARITHMETIC: four bilinear corners
  PM_LOCAL = 1,   // ARITHMETIC_LOCAL: nearest corner, its bilinear weight
  PM_HARM = 2,    // HARMONIC: nearest corner, weight 1, value 1/f
  PM_TIES = 3,    // subgridStressChanges.c:134-175 (<=0.5 / >=0.5, ties hit 2 or 4 nodes)
  PM_CENTRE = 4   // subgridStressChanges.c:223-233 (one cell-centre node)
};
enum { EPI_DIVIDE = 0, EPI_SUBGRID_T = 1, EPI_SUB_FROM = 2 };

#define FVM_MAXF 8
struct ProjSrc {
  const double *col;    // device column (marker index), or null
  double mat[FVM_MAXMAT];
  int use_mat;
};
struct ProjGroup {
  int stagx, stagy, method, nf;
  ProjSrc src[FVM_MAXF];
  double *out[FVM_MAXF];
  // epilogue: EPI_DIVIDE out = sum/w (HARM: w/sum)
  //           EPI_SUBGRID_T out = (a + -1*b) + -1*(sum/w)     (VecAXPBYPCZ)
  //           EPI_SUB_FROM  out = out + -1*(sum/w)             (VecAXPY(-1))
  int epilogue;
  const double *ea, *eb;
};

// findMarkerCells without the host round trip: counters (outside, bad) are
// left in ctx->red[0..1].  With `dens` it also runs the density loop of
// main_subduction.c:255-264; with build_order it also builds the bucket order
// for the new positions (sort.hip).
int fvm_locate_async(fvm_ctx *ctx, const fvm_grid *g, fvm_markers *mk, int32_t rank, const Materials *dens,
                     bool build_order);
// Build the (cell, quadrant) bucket order of the active markers (sort.hip).
// Bucket-ordered payload: position p of the sort gets marker perm[p]'s X, Y,
// Mat (with_xy) and the listed columns, so the partials passes read contiguous runs.
#define MAX_PAYLOAD_COLS 24
struct PayloadCols {
  int nc;
  const double *src[MAX_PAYLOAD_COLS];
  double *dst[MAX_PAYLOAD_COLS];
};
// Copy payload entry p from marker v.
__device__ __forceinline__ void gather_entry(const PayloadCols &pc, int with_xy, const double *X, const double *Y,
                                             const int32_t *Mat, double *xs, double *ys, int32_t *ms, int64_t p,
                                             int v) {
  if (with_xy) {
    xs[p] = X[v];
    ys[p] = Y[v];
    ms[p] = Mat[v];
  }
  for (int k = 0; k < pc.nc; k++) pc.dst[k][p] = pc.src[k][v];
}

// Put the four buckets of one basic cell (global bucket index 4c..4c+3) in
// ascending marker order and call emit(p, v) for every final position p and
// its marker v.  One wavefront per cell: the (usually <= 64) entries are ranked
// across lanes.  A fuller cell is ordered bucket by bucket: a bucket of up to
// 256 entries is held in registers (four per lane) and every entry is ranked
// against all of them (marker indices are distinct); only a bucket above 256
// entries falls back to an insertion sort by one lane (reading back its own
// writes), which population control keeps from happening.
template <class Emit>
__device__ __forceinline__ void bucket_order_cell(int64_t c, const int32_t *__restrict__ start,
                                                  int32_t *__restrict__ perm, Emit &&emit) {
  const int lane = threadIdx.x & 63;
  // c is uniform across the wavefront: make that explicit so that the bucket
  // starts and the loop bounds below live in scalar registers
  c = (int64_t)(((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(c >> 32)) << 32) |
                (uint32_t)__builtin_amdgcn_readfirstlane((int)c));
  const int b0 = start[4 * c], b1 = start[4 * c + 1], b2 = start[4 * c + 2], b3 = start[4 * c + 3];
  const int b4 = start[4 * c + 4];
  const int n = __builtin_amdgcn_readfirstlane(b4 - b0);
  if (n <= 64) {
    const int pos = b0 + lane;
    const bool have = lane < n;
    const int v = have ? perm[pos] : 0x7fffffff;
    const int bk = (pos >= b1) + (pos >= b2) + (pos >= b3);
    int p;
    if (__ballot(have && v >= (1 << 30)) == 0) {
      // (bucket, marker) packed in 32 bits: the final position is the cell's
      // start plus the number of the cell's entries with a smaller pair
      const uint32_t key = have ? ((uint32_t)bk << 30) | (uint32_t)v : 0xffffffffu;
      // lanes >= n hold the largest key, so the count may run to a multiple of 4
      int r0 = 0, r1 = 0, r2 = 0, r3 = 0;
      for (int j = 0; j < n; j += 4) {
        r0 += (uint32_t)read_lane((int)key, j) < key ? 1 : 0;
        r1 += (uint32_t)read_lane((int)key, j + 1) < key ? 1 : 0;
        r2 += (uint32_t)read_lane((int)key, j + 2) < key ? 1 : 0;
        r3 += (uint32_t)read_lane((int)key, j + 3) < key ? 1 : 0;
      }
      p = b0 + ((r0 + r1) + (r2 + r3));
    } else {
      int rank = 0;
      for (int j = 0; j < n; j++) {
        const int vj = read_lane(v, j);
        const int bj = read_lane(bk, j);
        rank += (bj == bk && vj < v) ? 1 : 0;
      }
      p = (bk == 0 ? b0 : (bk == 1 ? b1 : (bk == 2 ? b2 : b3))) + rank;
    }
    if (have) {
      perm[p] = v;
      emit(p, v);
    }
  } else {
#pragma unroll 1
    for (int q = 0; q < 4; q++) {
      const int s0 = q == 0 ? b0 : (q == 1 ? b1 : (q == 2 ? b2 : b3));
      const int s1 = q == 0 ? b1 : (q == 1 ? b2 : (q == 2 ? b3 : b4));
      const int len = s1 - s0;
      if (len <= 256) {
        int v[4], r[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
          const int pos = s0 + lane + 64 * u;
          v[u] = pos < s1 ? perm[pos] : 0x7fffffff;
          r[u] = 0;
        }
#pragma unroll
        for (int u2 = 0; u2 < 4; u2++) {
          if (64 * u2 >= len) break;
          for (int j = 0; j < 64; j++) {
            const int w = read_lane(v[u2], j);
#pragma unroll
            for (int u = 0; u < 4; u++) r[u] += w < v[u] ? 1 : 0;
          }
        }
        // every lane's loads completed before the ranks above: writing back
        // into the bucket's own range is safe
#pragma unroll
        for (int u = 0; u < 4; u++)
          if (lane + 64 * u < len) {
            perm[s0 + r[u]] = v[u];
            emit(s0 + r[u], v[u]);
          }
      } else if (lane == 0) {
        for (int i = s0 + 1; i < s1; i++) {
          const int32_t v = perm[i];
          int j = i - 1;
          while (j >= s0 && perm[j] > v) {
            perm[j + 1] = perm[j];
            j--;
          }
          perm[j + 1] = v;
        }
        for (int p = s0; p < s1; p++) emit(p, perm[p]);
      }
    }
  }
}

// Two basic cells per wavefront: c0 in lanes 0-31, c1 in lanes 32-63 (c1 < 0:
// none).  When both cells hold at most 32 entries (16-25 markers per cell are
// common) each half ranks its cell's packed (bucket, marker) keys against its
// own half, so one wavefront does the work of two; otherwise (or for marker
// indices >= 2^30) the cells go through bucket_order_cell one after the other.
// The order is the same either way: the cell's entries by (bucket, marker).
template <class Emit>
__device__ __forceinline__ void bucket_order_pair(int64_t c0, int64_t c1, const int32_t *__restrict__ start,
                                                  int32_t *__restrict__ perm, Emit &&emit) {
  c0 = (int64_t)(((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(c0 >> 32)) << 32) |
                 (uint32_t)__builtin_amdgcn_readfirstlane((int)c0));
  c1 = (int64_t)(((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(c1 >> 32)) << 32) |
                 (uint32_t)__builtin_amdgcn_readfirstlane((int)c1));
  if (c1 < 0) {
    if (c0 >= 0) bucket_order_cell(c0, start, perm, emit);
    return;
  }
  if (c0 < 0) {
    bucket_order_cell(c1, start, perm, emit);
    return;
  }
  const int a0 = start[4 * c0], a4 = start[4 * c0 + 4];
  const int d0 = start[4 * c1], d4 = start[4 * c1 + 4];
  const int n0 = __builtin_amdgcn_readfirstlane(a4 - a0), n1 = __builtin_amdgcn_readfirstlane(d4 - d0);
  if (n0 <= 32 && n1 <= 32) {
    const int lane = threadIdx.x & 63;
    const bool hi = lane >= 32;
    const int hl = lane & 31;
    const int64_t c = hi ? c1 : c0;
    const int b0 = start[4 * c], b1 = start[4 * c + 1], b2 = start[4 * c + 2], b3 = start[4 * c + 3];
    const int n = hi ? n1 : n0;
    const int pos = b0 + hl;
    const bool have = hl < n;
    const int v = have ? perm[pos] : 0x7fffffff;
    const int bk = (pos >= b1) + (pos >= b2) + (pos >= b3);
    if (__ballot(have && v >= (1 << 30)) == 0) {
      const uint32_t key = have ? ((uint32_t)bk << 30) | (uint32_t)v : 0xffffffffu;
      // each half counts the smaller keys of its own half: lanes j and 32 + j
      // are read together (j + 3 <= 31, so 32 + j + 3 stays a lane index)
      const int nmax = n0 > n1 ? n0 : n1;
      int r0 = 0, r1 = 0, r2 = 0, r3 = 0;
      for (int j = 0; j < nmax; j += 4) {
        const uint32_t l0 = (uint32_t)read_lane((int)key, j), h0 = (uint32_t)read_lane((int)key, 32 + j);
        const uint32_t l1 = (uint32_t)read_lane((int)key, j + 1), h1 = (uint32_t)read_lane((int)key, 33 + j);
        const uint32_t l2 = (uint32_t)read_lane((int)key, j + 2), h2 = (uint32_t)read_lane((int)key, 34 + j);
        const uint32_t l3 = (uint32_t)read_lane((int)key, j + 3), h3 = (uint32_t)read_lane((int)key, 35 + j);
        r0 += (hi ? h0 : l0) < key ? 1 : 0;
        r1 += (hi ? h1 : l1) < key ? 1 : 0;
        r2 += (hi ? h2 : l2) < key ? 1 : 0;
        r3 += (hi ? h3 : l3) < key ? 1 : 0;
      }
      const int p = b0 + ((r0 + r1) + (r2 + r3));
      if (have) {
        perm[p] = v;
        emit(p, v);
      }
      return;
    }
    // a marker index >= 2^30: nothing written yet, order the cells one by one
  }
  bucket_order_cell(c0, start, perm, emit);
  bucket_order_cell(c1, start, perm, emit);
}

struct SortGather {
  int with_xy;
  const double *X, *Y;
  const int32_t *Mat;
  double *xs, *ys;
  int32_t *ms;
  PayloadCols pc;
};
// Build (or reuse) the bucket order.  When the order has to be rebuilt and `sg`
// is given, the payload it names is gathered in the same pass and *gathered is
// set; a reused order leaves *gathered false.
int fvm_sort_markers(fvm_ctx *ctx, const fvm_grid *g, const fvm_markers *mk, const SortGather *sg = nullptr,
                     bool *gathered = nullptr);
bool fvm_sort_valid(const fvm_ctx *ctx, const fvm_grid *g, const fvm_markers *mk);
// Run projection groups over the current bucket order (project.hip).
int fvm_project_group(fvm_ctx *ctx, const fvm_grid *g, const fvm_markers *mk, const ProjGroup &pg);
struct SubT;
int fvm_project_groups(fvm_ctx *ctx, const fvm_grid *g, const fvm_markers *mk, int ng, const ProjGroup *pgs,
                       const SubT *sub = nullptr);
// Invalidate the cached bucket order (positions or cells changed).
static inline void fvm_markers_moved(fvm_ctx *ctx) { ctx->epoch++; }

// The columns of a marker store in a fixed order (context.hip): FVM_NUM_DCOLS
// FP64 columns (X first) and FVM_NUM_ICOLS int32 columns (Mat first).
#define FVM_NUM_DCOLS 23
#define FVM_NUM_ICOLS 5
void fvm_marker_columns(fvm_markers *mk, double **d, int32_t **i);

// ---- per-phase timing --------------------------------------------------------
struct ProfScope {
  fvm_ctx *ctx;
  int slot;
  ProfScope(fvm_ctx *c, int phase) : ctx(c), slot(-1) {
    if (c->profiling && c->nprof < 1024) {
      slot = c->nprof++;
      c->prof_phase[slot] = phase;
      (void)hipEventRecord(c->prof_ev[2 * slot], c->stream);
    }
  }
  ~ProfScope() {
    if (slot >= 0) (void)hipEventRecord(ctx->prof_ev[2 * slot + 1], ctx->stream);
  }
};
#define FVM_PHASE(ctx, ph) ProfScope fvm_prof_scope_(ctx, ph)

// subgridTemperatureChanges split for fusion (ops.hip)
int fvm_subgrid_t_pass1(fvm_ctx *ctx, const double *thisT, const fvm_grid *g, const NodalArrays *nf,
                        fvm_markers *mk, const Materials *mat, double dt, double dsubgrid, double **dTsub_out,
                        double **deltaT_out);
int fvm_advect_with_subgrid_t2(fvm_ctx *ctx, fvm_markers *mk, const NodalArrays *nf, const fvm_grid *g,
                               const BoundaryValues *bv, double dt, const double *deltaT, const double *dTsub,
                               double dt_thermal);

// Jacobi BiCGStab (pressure_guess.hip) on a 5-point system over the NX*NY
// nodes in node order: diagonal d, east (i+1), west (i-1), south (i+NX) and
// north (i-NX) coefficients, 0 where a row has no such neighbour; x0 = 0.
// Writes x, synchronises; FVM_ERR_NOT_CONVERGED when rtol is not reached.
int fvm_bicgstab_5pt(fvm_ctx *ctx, int NX, int NY, const double *d, const double *e, const double *w,
                     const double *sth, const double *nth, const double *b, double *x, double rtol, int32_t maxit,
                     int32_t *iters, double *relres);

// sort building blocks reused by the batched ensemble (sort.hip)
// One launch: count[0..ncount) = 0, counters[0..1] = 0, state[0..nstate) = 0.
int fvm_launch_zero_sort(fvm_ctx *ctx, int32_t *count, int64_t ncount, int32_t *counters, unsigned long long *state,
                         int64_t nstate);
// Exclusive scan of count[0..nk) into start[0..nk], start[nk] = total, in one
// pass; `state` holds fvm_scan_tiles(nk) + 2 zeroed words.
int fvm_launch_scan_onepass(fvm_ctx *ctx, const int32_t *count, int32_t *start, unsigned long long *state,
                            int64_t nk);
int64_t fvm_scan_tiles(int64_t nk);
__device__ __forceinline__ int bisect_grid(const double *__restrict__ x, int N, double X) {
  // findMarkerCell's bisection (markers.c:48-59): largest i in [0, N-2] with x[i] <= X
  int lo = 0, hi = N - 1;
  while (hi - lo > 1) {
    int mid = (hi + lo) / 2;
    if (mid == lo) mid++;
    else if (mid == hi) mid--;
    if (X >= x[mid]) lo = mid;
    else hi = mid;
  }
  return lo;
}

// The same index from a guess: for a non-decreasing x, bisect_grid returns the
// largest i in [0, N-2] with x[i] <= X (0 when there is none), whatever path
// the bisection takes.  Start at the cell a uniform spacing would give
// (inv = (N-1) / (x[N-1] - x[0])), step at most four cells towards the answer,
// and bisect when that did not settle it (strongly graded grids).
__device__ __forceinline__ int guess_grid(const double *__restrict__ x, int N, double X, double inv) {
  double t = (X - x[0]) * inv;
  t = t > 0.0 ? (t < (double)(N - 2) ? t : (double)(N - 2)) : 0.0;  // NaN -> 0
  int i = (int)t;
  bool done = false;
  for (int s = 0; s < 4 && !done; s++) {
    if (i > 0 && X < x[i]) i--;
    else if (i < N - 2 && X >= x[i + 1]) i++;
    else done = true;  // x[i] <= X < x[i+1], or clamped at an end
  }
  return done ? i : bisect_grid(x, N, X);
}

// deterministic double sum of n values on the device -> *out (host), ops.hip
int fvm_reduce_sum(fvm_ctx *ctx, const double *d, int64_t n, double *out);
