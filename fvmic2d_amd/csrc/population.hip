// Marker population control: checkMarkerDensity (markers.c:455-781) on one GPU.
//
// Quarter cells are the sort's buckets: bucket ((cellY*(NX-1)+cellX) << 2 | bx |
// by << 1) is the reference's quarter (i, j) = (2 cellX + bx, 2 cellY + by), and
// its members are listed in ascending marker index by the bucket order.  The
// reference's loops become:
//   plan     one thread per quarter (reference order q = i + 2(NX-1) j): target
//            and change, then exclusive scans of the additions and removals;
//   out list the markers without a cell, in index order (flag + scan);
//   remove   one thread per over-full quarter: its geometric members in index
//            order get keys from the counter stream, the lowest leave (appended to
//            the out list in that order);
//   new      one thread per added marker: its quarter by binary search over the
//            addition offsets, its position, its slot (recycled from the end of
//            the out list, else appended), and for an ordinary marker the nearest
//            original marker inside the one-cell window, searched over the
//            buckets of the cells that window can reach;
//   chain    one workgroup, in the reference's order: recycled slots filled
//            earlier in the call are candidates as well (only when slots are
//            recycled), and every added marker ends up naming the marker whose
//            fields it copies (an original marker or a slab / inflow marker);
//   write    one thread per added marker: copy (or build) its fields.
// Random numbers come from the counter stream of fvm_abi.h (SplitMix64 keyed by
// seed and stream), so the result is a pure function of the inputs and the seed.
#include <math.h>

#include <algorithm>

#include "fvm_internal.h"
#include "rheology.h"

namespace {

__device__ __forceinline__ uint64_t pc_mix(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__device__ __forceinline__ double pc_u01(uint64_t seed, uint64_t stream, uint64_t ctr) {
  const uint64_t key = pc_mix(seed + 0x9E3779B97F4A7C15ull * (stream + 1));
  return (double)(pc_mix(key ^ ctr) >> 11) * (1.0 / 9007199254740992.0);
}

struct PcParams {
  DGrid g;
  int32_t W;        // 2 (NX - 1): quarters per row
  int64_t nq;       // 4 (NX - 1)(NY - 1)
  int32_t nMark;    // markers before the call
  int32_t minMarkers, maxMarkers, NMX, NMY;
  uint64_t seed;
  double plateDepth;  // plate_depth(LX) (benchmarkInitialConditions.c:80-92)
  double slabAngleTan, slabDenom;  // tan(slabAngle), 2 sqrt(kappa * slab age in s)
  double wmax, hmax;  // largest cell width / height (nearest-neighbour window)
};

// quarter q (reference order) -> its bucket in the sort
__device__ __forceinline__ int64_t q_bucket(int64_t q, int W, int ncx) {
  const int i = (int)(q % W), j = (int)(q / W);
  return ((int64_t)((j >> 1) * ncx + (i >> 1)) << 2) | (i & 1) | ((j & 1) << 1);
}

// markers.c:583-601 / 670-679: bounds of quarter column i (row j likewise)
__device__ __forceinline__ void q_bounds(const double *x, const double *xc, int i, double &lo, double &hi) {
  if (!(i % 2)) {
    lo = x[i / 2];
    hi = xc[i / 2 + 1];
  } else {
    lo = xc[(i + 1) / 2];
    hi = x[(i + 1) / 2];
  }
}

__device__ __forceinline__ bool in_quarter(double X, double Y, double xmin, double xmax, double ymin, double ymax) {
  return X >= xmin && X < xmax && Y >= ymin && Y < ymax;
}

// markers.c:520-562: target of an under-full quarter, or the excess of an over-full
// one.  add/rem are the changes; geo the geometric members of an over-full quarter.
__global__ void __launch_bounds__(256) k_pc_plan(PcParams P, const int32_t *__restrict__ start,
                                                 const int32_t *__restrict__ perm, const double *__restrict__ X,
                                                 const double *__restrict__ Y, int32_t *__restrict__ add,
                                                 int32_t *__restrict__ rem, int32_t *__restrict__ flags) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= P.nq) return;
  const int64_t b = q_bucket(q, P.W, P.g.NX - 1);
  const int32_t c0 = start[b], c1 = start[b + 1];
  const int cnt = c1 - c0;
  const int i = (int)(q % P.W), j = (int)(q / P.W);
  int a = 0, r = 0;
  if (cnt < P.minMarkers) {
    double dy = P.g.y[j / 2 + 1] - P.g.y[j / 2];
    double dx = P.g.x[i / 2 + 1] - P.g.x[i / 2];
    dx /= 2.0;
    dy /= 2.0;
    const int mynmx = (int)ceil(P.NMX * (P.g.NX - 1) / 2.0 * dx / P.g.LX);
    const int mynmy = (int)ceil(P.NMY * (P.g.NY - 1) / 2.0 * dy / P.g.LY);
    a = mynmx * mynmy - cnt;
    if (a <= 0) {
      __hip_atomic_store(&flags[0], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      a = 0;
    }
  } else if (cnt > P.maxMarkers) {
    // only geometric members can leave (members on x = LX or y = LY are counted
    // by their cell but lie outside every quarter's half-open bounds)
    double xmin, xmax, ymin, ymax;
    q_bounds(P.g.x, P.g.xc, i, xmin, xmax);
    q_bounds(P.g.y, P.g.yc, j, ymin, ymax);
    int geo = 0;
    for (int p = c0; p < c1; p++) {
      const int m = perm[p];
      geo += in_quarter(X[m], Y[m], xmin, xmax, ymin, ymax) ? 1 : 0;
    }
    r = min(cnt - P.maxMarkers, geo);
  }
  add[q] = a;
  rem[q] = r;
}

__global__ void __launch_bounds__(256) k_pc_out_flags(int32_t n, const int32_t *__restrict__ cellX,
                                                      int32_t *__restrict__ flag) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m < n) flag[m] = cellX[m] == -1 ? 1 : 0;
}

__global__ void __launch_bounds__(256) k_pc_out_list(int32_t n, const int32_t *__restrict__ flag,
                                                     const int32_t *__restrict__ pos, int32_t *__restrict__ isout) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m < n && flag[m]) isout[pos[m]] = (int32_t)m;
}

// markers.c:616-663: the rem[q] members with the lowest (key, position) leave,
// appended to the out list in that order; keys u(seed, 1, q << 32 | k) for the
// k-th geometric member.  Positions and cells of the leavers are written after
// the selection, which reads them.
__global__ void __launch_bounds__(256) k_pc_remove(PcParams P, const int32_t *__restrict__ start,
                                                   const int32_t *__restrict__ perm, const int32_t *__restrict__ rem,
                                                   const int32_t *__restrict__ remOff, int32_t nOut0,
                                                   double *__restrict__ X, double *__restrict__ Y,
                                                   int32_t *__restrict__ cellX, int32_t *__restrict__ cellY,
                                                   int32_t *__restrict__ isout) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= P.nq) return;
  const int nr = rem[q];
  if (nr <= 0) return;
  const int64_t b = q_bucket(q, P.W, P.g.NX - 1);
  const int32_t c0 = start[b], c1 = start[b + 1];
  const int i = (int)(q % P.W), j = (int)(q / P.W);
  double xmin, xmax, ymin, ymax;
  q_bounds(P.g.x, P.g.xc, i, xmin, xmax);
  q_bounds(P.g.y, P.g.yc, j, ymin, ymax);
  // the r-th leaver is the smallest (key, k) above the (r-1)-th
  double pu = -1.0;
  int pk = -1;
  int32_t *dst = isout + nOut0 + remOff[q];
  for (int r = 0; r < nr; r++) {
    double bu = 2.0;
    int bk = -1, bm = -1;
    int k = 0;
    for (int p = c0; p < c1; p++) {
      const int m = perm[p];
      if (!in_quarter(X[m], Y[m], xmin, xmax, ymin, ymax)) continue;
      const double u = pc_u01(P.seed, 1, ((uint64_t)q << 32) | (uint64_t)k);
      const bool after = u > pu || (u == pu && k > pk);
      if (after && (u < bu || (u == bu && k < bk))) {
        bu = u;
        bk = k;
        bm = m;
      }
      k++;
    }
    dst[r] = bm;
    pu = bu;
    pk = bk;
  }
  // every member at or below the last pick leaves
  int k = 0;
  for (int p = c0; p < c1; p++) {
    const int m = perm[p];
    if (!in_quarter(X[m], Y[m], xmin, xmax, ymin, ymax)) continue;
    const double u = pc_u01(P.seed, 1, ((uint64_t)q << 32) | (uint64_t)k);
    if (u < pu || (u == pu && k <= pk)) {
      X[m] = -1.0;
      Y[m] = -1.0;
      cellX[m] = -1;
      cellY[m] = -1;
    }
    k++;
  }
}

// kinds of added marker (markers.c:713-719)
#define PC_NEAREST 0
#define PC_SLAB 1
#define PC_INFLOW 2

struct NewMarkers {
  double *X, *Y, *dist;
  int32_t *slot, *kind, *best, *cx, *cy, *valid, *root;
};

// The first candidate loop of markers.c:722-738 over the original markers: the
// nearest marker that still has a cell, within one width / height of its own
// cell; (distance, index) decides, as the reference's ascending scan with a
// strict '<' does.
__device__ __forceinline__ void nearest_original(const PcParams &P, double nx, double ny, int32_t self,
                                                 const int32_t *__restrict__ start, const int32_t *__restrict__ perm,
                                                 const double *__restrict__ X, const double *__restrict__ Y,
                                                 const int32_t *__restrict__ cellX, const int32_t *__restrict__ cellY,
                                                 double &bd, int32_t &bm) {
  const DGrid &g = P.g;
  // any candidate lies within [n - wmax, n + wmax]: scan the cells meeting it
  const double xa = fmax(nx - P.wmax, 0.0), xb = fmin(nx + P.wmax, g.LX);
  const double ya = fmax(ny - P.hmax, 0.0), yb = fmin(ny + P.hmax, g.LY);
  const int cx0 = bisect_grid(g.x, g.NX, xa), cx1 = bisect_grid(g.x, g.NX, xb);
  const int cy0 = bisect_grid(g.y, g.NY, ya), cy1 = bisect_grid(g.y, g.NY, yb);
  bd = g.LX * g.LY;  // markers.c:722: nothing at or beyond LX*LY is taken
  bm = -1;
  for (int cy = cy0; cy <= cy1; cy++) {
    for (int cx = cx0; cx <= cx1; cx++) {
      const int64_t b = (int64_t)(cy * (g.NX - 1) + cx) << 2;
      const int32_t p0 = start[b], p1 = start[b + 4];
      for (int p = p0; p < p1; p++) {
        const int32_t m1 = perm[p];
        if (m1 == self || m1 >= P.nMark) continue;
        const int c = cellX[m1], r = cellY[m1];
        if (c == -1) continue;  // removed in this call
        if (fabs(X[m1] - nx) <= (g.x[c + 1] - g.x[c]) && fabs(Y[m1] - ny) <= (g.y[r + 1] - g.y[r])) {
          const double x1 = (X[m1] - nx);
          const double y1 = (Y[m1] - ny);
          const double r1 = sqrt(x1 * x1 + y1 * y1);
          if (r1 < bd || (r1 == bd && m1 < bm)) {
            bd = r1;
            bm = m1;
          }
        }
      }
    }
  }
}

// markers.c:684-711 and 713-719: position, slot and kind of every added marker,
// and its nearest original neighbour
__global__ void __launch_bounds__(256) k_pc_new(PcParams P, int32_t nAdd, const int32_t *__restrict__ addOff,
                                                const int32_t *__restrict__ isout, int32_t nOut,
                                                const int32_t *__restrict__ start, const int32_t *__restrict__ perm,
                                                const double *__restrict__ X, const double *__restrict__ Y,
                                                const int32_t *__restrict__ cellX, const int32_t *__restrict__ cellY,
                                                NewMarkers nm) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nAdd) return;
  // quarter: the last q with addOff[q] <= i
  int64_t lo = 0, hi = P.nq;  // addOff[lo] <= i < addOff[hi]
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) >> 1;
    if (addOff[mid] <= i) lo = mid;
    else hi = mid;
  }
  const int qi = (int)(lo % P.W), qj = (int)(lo / P.W);
  double a, b;
  q_bounds(P.g.x, P.g.xc, qi, a, b);
  const double nx = (b - a) * pc_u01(P.seed, 2, (uint64_t)i) + a;
  q_bounds(P.g.y, P.g.yc, qj, a, b);
  const double ny = (b - a) * pc_u01(P.seed, 3, (uint64_t)i) + a;
  const int32_t slot = i < nOut ? isout[nOut - 1 - i] : P.nMark + (int32_t)(i - nOut);
  int kind = PC_NEAREST;
  if (nx < P.g.x[1]) kind = PC_SLAB;
  else if (nx >= P.g.x[P.g.NX - 2] && ny >= P.plateDepth) kind = PC_INFLOW;
  nm.X[i] = nx;
  nm.Y[i] = ny;
  nm.slot[i] = slot;
  nm.kind[i] = kind;
  if (kind == PC_NEAREST) {
    double bd;
    int32_t bm;
    nearest_original(P, nx, ny, slot, start, perm, X, Y, cellX, cellY, bd, bm);
    nm.dist[i] = bd;
    nm.best[i] = bm;
    nm.valid[i] = bm >= 0;
    nm.root[i] = bm;
    nm.cx[i] = bm >= 0 ? cellX[bm] : -1;
    nm.cy[i] = bm >= 0 ? cellY[bm] : -1;
  } else {
    // findMarkerCell of the new position (inside the domain)
    nm.cx[i] = bisect_grid(P.g.x, P.g.NX, nx);
    nm.cy[i] = bisect_grid(P.g.y, P.g.NY, ny);
    nm.valid[i] = 1;
    nm.root[i] = -1 - (int32_t)i;  // builds its own fields
    nm.best[i] = -1;
    nm.dist[i] = P.g.LX * P.g.LY;
  }
}

// Recycled slots filled earlier in the call join the candidates of the later
// markers (markers.c:722-738 reads the array as it is being refilled).  One
// workgroup walks the added markers in order; for each, its threads scan the
// earlier recycled ones.
#define CHAIN_BS 256
__global__ void __launch_bounds__(CHAIN_BS) k_pc_chain(PcParams P, int32_t nAdd, int32_t nOut, NewMarkers nm) {
  __shared__ double s_d[CHAIN_BS];
  __shared__ int32_t s_m[CHAIN_BS], s_j[CHAIN_BS];
  const DGrid &g = P.g;
  for (int32_t i = 0; i < nAdd; i++) {
    if (nm.kind[i] != PC_NEAREST) continue;  // uniform over the block
    const double nx = nm.X[i], ny = nm.Y[i];
    double bd = g.LX * g.LY;  // as the reference's starting distance
    int32_t bm = -1, bj = -1;
    const int32_t jend = min(i, nOut);
    for (int32_t j = threadIdx.x; j < jend; j += CHAIN_BS) {
      if (!nm.valid[j]) continue;
      const int c = nm.cx[j], r = nm.cy[j];
      const double xj = nm.X[j], yj = nm.Y[j];
      if (fabs(xj - nx) <= (g.x[c + 1] - g.x[c]) && fabs(yj - ny) <= (g.y[r + 1] - g.y[r])) {
        const double x1 = (xj - nx);
        const double y1 = (yj - ny);
        const double r1 = sqrt(x1 * x1 + y1 * y1);
        const int32_t m1 = nm.slot[j];
        if (r1 < bd || (r1 == bd && m1 < bm)) {
          bd = r1;
          bm = m1;
          bj = j;
        }
      }
    }
    s_d[threadIdx.x] = bd;
    s_m[threadIdx.x] = bm;
    s_j[threadIdx.x] = bj;
    __syncthreads();
    for (int o = CHAIN_BS / 2; o > 0; o >>= 1) {
      if ((int)threadIdx.x < o) {
        const double d2 = s_d[threadIdx.x + o];
        const int32_t m2 = s_m[threadIdx.x + o];
        if (m2 >= 0 && (s_m[threadIdx.x] < 0 || d2 < s_d[threadIdx.x] ||
                        (d2 == s_d[threadIdx.x] && m2 < s_m[threadIdx.x]))) {
          s_d[threadIdx.x] = d2;
          s_m[threadIdx.x] = m2;
          s_j[threadIdx.x] = s_j[threadIdx.x + o];
        }
      }
      __syncthreads();
    }
    if (threadIdx.x == 0 && s_m[0] >= 0) {
      const double d0 = nm.dist[i];
      const int32_t m0 = nm.best[i];
      if (m0 < 0 || s_d[0] < d0 || (s_d[0] == d0 && s_m[0] < m0)) {
        const int32_t j = s_j[0];
        nm.root[i] = nm.root[j];
        nm.cx[i] = nm.cx[j];
        nm.cy[i] = nm.cy[j];
        nm.valid[i] = 1;
      }
    }
    __syncthreads();
  }
}

struct MarkerCols {
  double *d[23];
  int32_t *n[5];  // Mat, cellX, cellY, cpu, isYielding
};

// benchmarkInitialConditions.c:52-72 on a resetMarker'd marker (markers.c:255-330;
// cpu and strainRateResidual, left unset by the reference, are 0); the columns
// follow fvm_markers' order: X Y VX VY T Tdot sxx syy sxy eta mu rho rhodot p exx
// eyy exy Exx Eyy Exy Eii wxy strainRateResidual
__device__ __forceinline__ int build_special(const PcParams &P, const YieldParams &Y, int kind, double x, double y,
                                            double *d, int32_t *n) {
  for (int k = 0; k < 23; k++) d[k] = 0.0;
  d[0] = x;
  d[1] = y;
  n[0] = 0;  // Mat
  n[3] = 0;  // cpu
  n[4] = 0;  // isYielding
  double T;
  if (kind == PC_SLAB) {
    const double Ts = 273, T0 = 1623;
    T = Ts + (T0 - Ts) * erf((y - x * P.slabAngleTan) / P.slabDenom);
  } else {
    T = 1573.0;
  }
  d[4] = T;
  double eta = 0.0, mu = 0.0;
  const int bad = rheology(Y, 0, T, x, y, &eta, &mu);
  d[9] = eta;
  d[10] = mu;
  return bad;
}

__global__ void __launch_bounds__(256) k_pc_write(PcParams P, YieldParams Yp, int32_t nAdd, NewMarkers nm,
                                                  MarkerCols mc, int32_t *__restrict__ flags) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nAdd || !nm.valid[i]) return;  // no neighbour: the slot is left as it is
  const int32_t slot = nm.slot[i], root = nm.root[i];
  double d[23];
  int32_t n[5];
  if (root >= 0) {
#pragma unroll
    for (int k = 0; k < 23; k++) d[k] = mc.d[k][root];
#pragma unroll
    for (int k = 0; k < 5; k++) n[k] = mc.n[k][root];
  } else {
    const int32_t k0 = -1 - root;  // the slab / inflow marker whose fields are copied
    if (build_special(P, Yp, nm.kind[k0], nm.X[k0], nm.Y[k0], d, n))
      __hip_atomic_store(&flags[1], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    n[1] = nm.cx[k0];
    n[2] = nm.cy[k0];
  }
  d[0] = nm.X[i];
  d[1] = nm.Y[i];
#pragma unroll
  for (int k = 0; k < 23; k++) mc.d[k][slot] = d[k];
#pragma unroll
  for (int k = 0; k < 5; k++) mc.n[k][slot] = n[k];
}

__global__ void k_pc_zero(int32_t *__restrict__ a, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0;
}

// largest cell width and height (one block)
__global__ void __launch_bounds__(256) k_pc_spacing(DGrid g, double *out) {
  __shared__ double s_w[256], s_h[256];
  double w = 0.0, h = 0.0;
  for (int i = threadIdx.x; i < g.NX - 1; i += 256) w = fmax(w, g.x[i + 1] - g.x[i]);
  for (int i = threadIdx.x; i < g.NY - 1; i += 256) h = fmax(h, g.y[i + 1] - g.y[i]);
  s_w[threadIdx.x] = w;
  s_h[threadIdx.x] = h;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      s_w[threadIdx.x] = fmax(s_w[threadIdx.x], s_w[threadIdx.x + o]);
      s_h[threadIdx.x] = fmax(s_h[threadIdx.x], s_h[threadIdx.x + o]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = s_w[0];
    out[1] = s_h[0];
  }
}

int64_t scan_state_words(int64_t n) { return fvm_scan_tiles(n) + 2; }

}  // namespace

// markers.c:455-781
extern "C" int fvm_check_marker_density(fvm_ctx *ctx, const fvm_grid *g, fvm_markers *mk, const Materials *mat,
                                        const Options *opt, uint64_t seed, int32_t *nAdded, int32_t *nRemoved) {
  if (!ctx || !mk || !mat || !opt) return FVM_ERR_ARG;
  FVM_TRY(fvm_grid_check_nonperiodic(ctx, g, "checkMarkerDensity"));
  if (nAdded) *nAdded = 0;
  if (nRemoved) *nRemoved = 0;
  const int32_t n = mk->nMark;
  const int64_t nq = 4LL * (g->NX - 1) * (g->NY - 1);
  // the bucket order of the current cells (the caller ran findMarkerCells)
  FVM_TRY(fvm_sort_markers(ctx, g, mk));
  const int32_t *start = (const int32_t *)ctx->start.ptr, *perm = (const int32_t *)ctx->perm.ptr;
  // phase-1 workspace: add, rem (nq), addOff, remOff (nq + 1), out flags / positions
  // (n, n + 1), out list (n), two scan states, spacing (2 doubles), flags
  const int64_t nn = n > 0 ? n : 1;
  const int64_t sq = scan_state_words(nq), sn = scan_state_words(nn);
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t o = off;
    off += (bytes + 255) & ~(size_t)255;
    return o;
  };
  const size_t o_add = take(nq * 4), o_rem = take(nq * 4), o_addOff = take((nq + 1) * 4),
               o_remOff = take((nq + 1) * 4), o_flag = take(nn * 4), o_pos = take((nn + 1) * 4),
               o_out = take(nn * 4), o_s1 = take(sq * 8), o_s2 = take(sq * 8), o_s3 = take(sn * 8),
               o_sp = take(16), o_fl = take(16);
  FVM_TRY(fvm_ws(ctx, ctx->popc, off));
  char *base = (char *)ctx->popc.ptr;
  int32_t *add = (int32_t *)(base + o_add), *rem = (int32_t *)(base + o_rem);
  int32_t *addOff = (int32_t *)(base + o_addOff), *remOff = (int32_t *)(base + o_remOff);
  int32_t *flag = (int32_t *)(base + o_flag), *pos = (int32_t *)(base + o_pos), *isout = (int32_t *)(base + o_out);
  unsigned long long *s1 = (unsigned long long *)(base + o_s1), *s2 = (unsigned long long *)(base + o_s2),
                     *s3 = (unsigned long long *)(base + o_s3);
  double *spacing = (double *)(base + o_sp);
  int32_t *flags = (int32_t *)(base + o_fl);

  PcParams P;
  P.g = dgrid(g);
  P.W = 2 * (g->NX - 1);
  P.nq = nq;
  P.nMark = n;
  P.minMarkers = opt->minMarkers;
  P.maxMarkers = opt->maxMarkers;
  P.NMX = opt->NMX;
  P.NMY = opt->NMY;
  P.seed = seed;
  {  // plate_depth(LX) and the slab constants, with the host C library
    const double x = g->LX;
    P.plateDepth = fabs(x - opt->rootCenter) < opt->rootWidth
                       ? opt->overridingPlateThickness +
                             opt->rootThickness * pow(cos(M_PI * (x - opt->rootCenter) / opt->rootWidth / 2.0), 2.0)
                       : opt->overridingPlateThickness;
    // slab_inflow_temperature's constants (benchmarkInitialConditions.c:124-131)
    const double kappa = 0.7272e-6, s_in_year = 3.15576e7;
    const double slab_age_s = opt->slabAge * 1.0e6 * s_in_year;
    P.slabAngleTan = tan(opt->slabAngle);
    P.slabDenom = 2.0 * sqrt(kappa * slab_age_s);
  }
  FVM_PHASE(ctx, FVM_PHASE_OTHER);
  // zero the scan states and the flags (kernels, no memset nodes)
  {
    int32_t *zs[4] = {(int32_t *)s1, (int32_t *)s2, (int32_t *)s3, flags};
    const int64_t zn[4] = {2 * sq, 2 * sq, 2 * sn, 4};
    for (int k = 0; k < 4; k++) {
      k_pc_zero<<<fvm_blocks(zn[k], 256), 256, 0, ctx->stream>>>(zs[k], zn[k]);
      FVM_LAUNCH_CHECK(ctx);
    }
  }
  k_pc_spacing<<<1, 256, 0, ctx->stream>>>(P.g, spacing);
  FVM_LAUNCH_CHECK(ctx);
  k_pc_plan<<<fvm_blocks(nq, 256), 256, 0, ctx->stream>>>(P, start, perm, mk->X, mk->Y, add, rem, flags);
  FVM_LAUNCH_CHECK(ctx);
  FVM_TRY(fvm_launch_scan_onepass(ctx, add, addOff, s1, nq));
  FVM_TRY(fvm_launch_scan_onepass(ctx, rem, remOff, s2, nq));
  if (n > 0) {
    k_pc_out_flags<<<fvm_blocks(n, 256), 256, 0, ctx->stream>>>(n, mk->cellX, flag);
    FVM_LAUNCH_CHECK(ctx);
    FVM_TRY(fvm_launch_scan_onepass(ctx, flag, pos, s3, n));
    k_pc_out_list<<<fvm_blocks(n, 256), 256, 0, ctx->stream>>>(n, flag, pos, isout);
    FVM_LAUNCH_CHECK(ctx);
  }
  // totals to the host: additions, removals, markers without a cell, flags
  int32_t *h = (int32_t *)ctx->pinned.ptr + 64;
  FVM_HIP(ctx, hipMemcpyAsync(h, addOff + nq, 4, hipMemcpyDeviceToHost, ctx->stream));
  FVM_HIP(ctx, hipMemcpyAsync(h + 1, remOff + nq, 4, hipMemcpyDeviceToHost, ctx->stream));
  if (n > 0) FVM_HIP(ctx, hipMemcpyAsync(h + 2, pos + n, 4, hipMemcpyDeviceToHost, ctx->stream));
  FVM_HIP(ctx, hipMemcpyAsync(h + 3, flags, 4, hipMemcpyDeviceToHost, ctx->stream));
  double *hd = (double *)ctx->pinned.ptr + 40;
  FVM_HIP(ctx, hipMemcpyAsync(hd, spacing, 16, hipMemcpyDeviceToHost, ctx->stream));
  FVM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  const int32_t nAdd = h[0], nRem = h[1], nOut0 = n > 0 ? h[2] : 0;
  if (h[3])
    return fvm_set_err(ctx, FVM_ERR_NEGATIVE_ADD,
                       "checkMarkerDensity: an under-full quarter's target does not exceed its count");
  P.wmax = hd[0];
  P.hmax = hd[1];
  if (nRem > 0) {
    k_pc_remove<<<fvm_blocks(nq, 256), 256, 0, ctx->stream>>>(P, start, perm, rem, remOff, nOut0, mk->X, mk->Y,
                                                              mk->cellX, mk->cellY, isout);
    FVM_LAUNCH_CHECK(ctx);
    fvm_markers_moved(ctx);
  }
  int32_t nOut = nOut0 + nRem;
  if (nRemoved) *nRemoved = nRem;
  const int64_t after = (int64_t)nAdd + n - nOut;
  if (nAdd > 0 && after < mk->maxMark) {
    // phase 2: the added markers
    size_t off2 = 0;
    auto take2 = [&](size_t bytes) {
      const size_t o = off2;
      off2 += (bytes + 255) & ~(size_t)255;
      return o;
    };
    const size_t a = (size_t)nAdd;
    const size_t oX = take2(a * 8), oY = take2(a * 8), oD = take2(a * 8), oS = take2(a * 4), oK = take2(a * 4),
                 oB = take2(a * 4), oCx = take2(a * 4), oCy = take2(a * 4), oV = take2(a * 4), oR = take2(a * 4);
    FVM_TRY(fvm_ws(ctx, ctx->popc2, off2));
    char *b2 = (char *)ctx->popc2.ptr;
    NewMarkers nm;
    nm.X = (double *)(b2 + oX);
    nm.Y = (double *)(b2 + oY);
    nm.dist = (double *)(b2 + oD);
    nm.slot = (int32_t *)(b2 + oS);
    nm.kind = (int32_t *)(b2 + oK);
    nm.best = (int32_t *)(b2 + oB);
    nm.cx = (int32_t *)(b2 + oCx);
    nm.cy = (int32_t *)(b2 + oCy);
    nm.valid = (int32_t *)(b2 + oV);
    nm.root = (int32_t *)(b2 + oR);
    k_pc_new<<<fvm_blocks(nAdd, 256), 256, 0, ctx->stream>>>(P, nAdd, addOff, isout, nOut, start, perm, mk->X,
                                                             mk->Y, mk->cellX, mk->cellY, nm);
    FVM_LAUNCH_CHECK(ctx);
    if (nOut > 0) {
      k_pc_chain<<<1, CHAIN_BS, 0, ctx->stream>>>(P, nAdd, nOut, nm);
      FVM_LAUNCH_CHECK(ctx);
    }
    MarkerCols mc;
    double *dc[23] = {mk->X,   mk->Y,   mk->VX,  mk->VY,  mk->T,   mk->Tdot, mk->sxx, mk->syy,
                      mk->sxy, mk->eta, mk->mu,  mk->rho, mk->rhodot, mk->p, mk->exx, mk->eyy,
                      mk->exy, mk->Exx, mk->Eyy, mk->Exy, mk->Eii, mk->wxy, mk->strainRateResidual};
    int32_t *ic[5] = {mk->Mat, mk->cellX, mk->cellY, mk->cpu, mk->isYielding};
    for (int k = 0; k < 23; k++) mc.d[k] = dc[k];
    for (int k = 0; k < 5; k++) mc.n[k] = ic[k];
    const YieldParams Yp = make_yield_params(mat, opt, 0.0);
    k_pc_write<<<fvm_blocks(nAdd, 256), 256, 0, ctx->stream>>>(P, Yp, nAdd, nm, mc, flags);
    FVM_LAUNCH_CHECK(ctx);
    FVM_HIP(ctx, hipMemcpyAsync(h + 3, flags + 1, 4, hipMemcpyDeviceToHost, ctx->stream));
    FVM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    fvm_markers_moved(ctx);
    if (nAdd > nOut) mk->nMark = n + (nAdd - nOut);
    if (nAdded) *nAdded = nAdd;
    if (h[3]) return fvm_set_err(ctx, FVM_ERR_VISCOSITY_LAW, "viscosity law undefined for material 0");
  } else if (after > mk->maxMark) {
    return fvm_set_err(ctx, FVM_ERR_CAPACITY, "checkMarkerDensity: %d markers to add, %d of %d slots free", nAdd,
                       mk->maxMark - (n - nOut), mk->maxMark);
  }
  return FVM_OK;
}
