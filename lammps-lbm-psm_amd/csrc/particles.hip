// Particle <-> lattice coupling kernels.
//
//  * prep/bin: one wavefront per list entry computes the reference's lattice
//    coordinates, radius, node-velocity terms and zones
//    (src/LBM_PSM_exchangeParticleData.cpp:31-55 and :175-195) and appends
//    the entry to the list of every node brick its map zone touches.
//  * map: one workgroup per touched brick.  The reference visits particles in
//    list order and, for each, the nodes of its zone; a node's slots depend
//    only on the ordered sequence of entries whose zone contains it.  Each
//    brick sorts its list by entry index and every node replays that
//    sequence, which reproduces the reference's slot state exactly
//    (src/LBM_PSM_exchangeParticleData.cpp:59-110,
//    src/LBM_PSM_lattice.cpp:257-324).
//  * hydro: one workgroup per entry sums the node forces of its slots into
//    force, torque and stresslet (src/LBM_PSM_exchangeParticleData.cpp:172-244)
//    with a fixed-order tree reduction.
//
// Compiled with -ffp-contract=off: all geometry is the reference's FP64.
#include <algorithm>

#include "lbmpsm_internal.h"
#include "solid_fraction.h"

namespace lbmpsm {

__device__ __forceinline__ int clampi(int v, int lo, int hi) {
  return v < lo ? lo : (v > hi ? hi : v);
}
__device__ __forceinline__ int maxi(int a, int b) { return a > b ? a : b; }
__device__ __forceinline__ int mini(int a, int b) { return a < b ? a : b; }

// Reference halo'd extent of an axis (nx+2, or 1 for z in 2D).
__device__ __forceinline__ int ref_extent(const Geom& g, int axis) {
  if (axis == 2 && g.dim == 2) return 1;
  const int n = axis == 0 ? g.nx : (axis == 1 ? g.ny : g.nz);
  return n + 2;
}
__device__ __forceinline__ int stored_extent(const Geom& g, int axis) {
  return axis == 0 ? g.sx : (axis == 1 ? g.sy : g.sz);
}

__global__ __launch_bounds__(256) void k_fill_int(int* p, int64_t n, int v) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t < n) p[t] = v;
}

__device__ __forceinline__ void brick_range(const int* z, int bi, int bj, int bk, int b[6]) {
  b[0] = z[0] / bi;
  b[1] = (z[1] - 1) / bi;
  b[2] = z[2] / bj;
  b[3] = (z[3] - 1) / bj;
  b[4] = z[4] / bk;
  b[5] = (z[5] - 1) / bk;
}

__device__ __forceinline__ bool zone_empty(const int* z) {
  return z[1] <= z[0] || z[3] <= z[2] || z[5] <= z[4];
}

// Reference lattice quantities of entry p; zones in stored coordinates.
__device__ __forceinline__ void entry_geometry(const PrepArgs& a, int p, double loc[3],
                                               double& r_lb, int z[6], int rz[6]) {
  const Geom& g = a.g;
  // src/LBM_PSM_exchangeParticleData.cpp:31-39 on one rank: the local shift
  // (nTotal - coord*2*envelope - 1) evaluates to -1.0 on every axis.
  const double off = (0 - ((double)0 * 2.0 * (double)1) - 1);
  for (int d = 0; d < 3; ++d) loc[d] = (a.x[3 * p + d] - a.origin[d]) / a.dx - off;
  r_lb = a.r[p] / a.dx;
  for (int d = 0; d < 3; ++d) {
    const int lo = (int)(loc[d] - r_lb) - 2;
    const int hi = (int)(loc[d] + r_lb) + 2;
    const int nh = ref_extent(g, d);
    // map zone, :44-48
    int mlo, mhi;
    if (d < 2) {
      mlo = clampi(lo, 0, nh);
      mhi = clampi(hi, 0, nh);
    } else {
      mlo = maxi(clampi(lo, 0, nh), 0);
      mhi = maxi(clampi(hi, 0, nh), 1);
    }
    // reduction zone, :191-195
    int rlo, rhi;
    if (d < 2) {
      rlo = clampi(lo, 1, nh - 1);
      rhi = clampi(hi, 1, nh - 1);
    } else {
      rlo = maxi(mini(maxi(lo, 1), nh - 1), 0);
      rhi = maxi(mini(maxi(hi, 1), nh - 1), 1);
    }
    const int se = stored_extent(g, d);
    z[2 * d] = clampi(mlo - g.off[d], 0, se);
    z[2 * d + 1] = clampi(mhi - g.off[d], 0, se);
    rz[2 * d] = clampi(rlo - g.off[d], 0, se);
    rz[2 * d + 1] = clampi(rhi - g.off[d], 0, se);
  }
}

// One wavefront per entry: every lane derives the same geometry, lane 0
// stores it, and the lanes share the appends to the touched bricks.
__global__ __launch_bounds__(64) void k_prep_bin(const PrepArgs a) {
  const int p = blockIdx.x;
  if (p == 0 && threadIdx.x == 0) *a.slow_count = 0;  // the map kernels of this step fill it
  if (p >= a.n) return;
  double loc[3], r_lb;
  int z[6], rz[6];
  entry_geometry(a, p, loc, r_lb, z, rz);
  if (threadIdx.x == 0) {
    if (a.tag_out) a.tag_out[p] = a.tag_in[p];
    a.xl[4 * p + 0] = loc[0];
    a.xl[4 * p + 1] = loc[1];
    a.xl[4 * p + 2] = loc[2];
    a.xl[4 * p + 3] = r_lb;
    for (int d = 0; d < 3; ++d) {
      a.kin[6 * p + d] = a.v[3 * p + d] / a.Uc * a.u_lb;      // get_vel_lb
      a.kin[6 * p + 3 + d] = a.w[3 * p + d] * a.tc * a.dt_d;  // get_freq_lb
    }
    for (int c = 0; c < 6; ++c) {
      a.zone[6 * p + c] = z[c];
      a.rzone[6 * p + c] = rz[c];
    }
  }
  if (zone_empty(z)) return;
  int b[6];
  brick_range(z, a.bi, a.bj, a.bk, b);
  const int nx = b[1] - b[0] + 1, ny = b[3] - b[2] + 1, nz = b[5] - b[4] + 1;
  const int tot = nx * ny * nz;
  for (int t = threadIdx.x; t < tot; t += 64) {
    const int x = b[0] + t / (ny * nz);
    const int y = b[2] + (t / nz) % ny;
    const int w = b[4] + t % nz;
    const int64_t id = ((int64_t)x * a.nbj + y) * a.nbk + w;
    const int slot = atomicAdd(&a.count[id], 1);
    if (slot < kBrickCap) a.list[id * kBrickCap + slot] = p;
  }
}

// Per-entry data the map kernel stages in LDS.
struct MapEntry {
  int zone[6];
  int tag;
  int raw;    // list index (staged lists: entries in list order)
  int coded;  // the zone has solid-fraction codes (zone_coded)
  double xl[4];
  double kin[6];
};

// z0 <= v < z1 for z0 <= z1, as one unsigned comparison
__device__ __forceinline__ bool in_range(int v, int z0, int z1) {
  return (unsigned)(v - z0) < (unsigned)(z1 - z0);
}

// One state-machine transition of a node for one entry whose zone contains
// it, given the entry's solid fraction at the node,
// src/LBM_PSM_exchangeParticleData.cpp:69-106.
struct NodeSlots {
  int t0, t1;
  double sf0, sf1;  // values of the slots written this step
  bool d0, d1;      // slot written this step
};

// The reference's decision (src/LBM_PSM_exchangeParticleData.cpp:69-106)
// is decFlag = (sf > 1e-5) + 2 (sf_old > 1e-5), sf_old being the fraction in
// this tag's slot (0 when the tag holds none): 0 clear, 1 set, 2 clear if
// the tag holds the slot, 3 set if sf > sf_old or the tag holds it.  A slot
// is only ever set with sf > 1e-5 and cleared to tag 0, sf 0, so sf_old >
// 1e-5 exactly when the tag holds a slot, and then cases 2 and 3 always act.
// The table is therefore: sf > 1e-5 sets the tag's slot, anything else
// clears it (a no-op when the tag holds none).  A tag 0 matching an empty
// slot reads sf_old = 0 and takes cases 0 / 1, which act the same.  The old
// fractions are never needed.
// Stores a node vector (three components, plane apart) at element o of the
// double array, or of its float view when the storage keeps node vectors in
// float.
__device__ __forceinline__ void store_node_vec(double* d, float* f, int64_t o, int64_t plane,
                                               double x, double y, double z) {
  if (f) {
    f[o] = (float)x;
    f[o + plane] = (float)y;
    f[o + 2 * plane] = (float)z;
  } else {
    d[o] = x;
    d[o + plane] = y;
    d[o + 2 * plane] = z;
  }
}

__device__ __forceinline__ void map_transition(const MapArgs& a, const MapEntry& E, double sf,
                                               int ri, int rj, int rk, int n, NodeSlots& S) {
  const int64_t plane = a.g.plane;
  const int tag = E.tag;
  const bool do_set = sf > 0.00001;
  const bool do_clear = !do_set;
  if (do_set) {  // setParticleOnLattice, src/LBM_PSM_lattice.cpp:257-299
    const double dx = (double)ri - E.xl[0];
    const double dy = (double)rj - E.xl[1];
    const double dz = (double)rk - E.xl[2];
    const double* kn = E.kin;
    double un[3];
    un[0] = kn[0] + kn[4] * dz - kn[5] * dy;
    un[1] = kn[1] + kn[5] * dx - kn[3] * dz;
    un[2] = kn[2] + kn[3] * dy - kn[4] * dx;
    int sl = -1;
    if (S.t0 == tag) sl = 0;
    else if (S.t1 == tag) sl = 1;
    else if (S.t0 == 0) sl = 0;
    else if (S.t1 == 0) sl = 1;
    // the node velocity is stored at once (a later entry of the node stores
    // after it); tag, solid fraction and force are stored at the end
    if (sl == 0) {
      S.t0 = tag; S.sf0 = sf; S.d0 = true;
    } else if (sl == 1) {
      S.t1 = tag; S.sf1 = sf; S.d1 = true;
    }
    if (sl >= 0) store_node_vec(a.s.us, a.s.usf, 3 * plane * sl + n, plane, un[0], un[1], un[2]);
  } else if (do_clear) {  // setToZero, src/LBM_PSM_lattice.cpp:302-324
    int sl = -1;
    if (S.t0 == tag) {
      S.t0 = 0; S.sf0 = 0.0; S.d0 = true;
      sl = 0;
    } else if (S.t1 == tag) {
      S.t1 = 0; S.sf1 = 0.0; S.d1 = true;
      sl = 1;
    }
    if (sl >= 0) store_node_vec(a.s.us, a.s.usf, 3 * plane * sl + n, plane, 0.0, 0.0, 0.0);
  }
}

// Solid fraction of one entry over its map zone.  Each workgroup takes one
// chunk of kSfChunk consecutive zone nodes (grid = entries x chunks per
// entry; chunks past the end of a zone exit):
//  1. every thread classifies its nodes with the reference's early exits;
//     cells that need the sub-point count are queued by column axis
//     (sf_column_axis);
//  2. the queued cells are counted in batches of kSfBatch cells, ten lanes
//     per cell, one lane per slice of the lane axis: each lane settles the
//     columns whose end points agree and queues the rest; the undecided
//     columns are then counted one per lane (ten uniform tests each).
// Every path sums the reference's own integer count (sf_count_columns).
// Codes are stored k fastest over the zone box (index within the zone).
constexpr int kSfChunk = 1024;
constexpr int kLineMax = 32;  // y/z zone extent of the line kernel (k_entry_sf_lines)
constexpr int kSfBatch = 24;
constexpr int kSfUndecCap = kSfBatch * kSfPairs;
static_assert(kSfBatch * kSfSlices <= 256, "one lane per cell slice");

// Zones whose codes k_entry_sf stores (others are computed inline by k_map):
// at most zone_cap nodes, at most zone_span per axis (the bound the code
// tables and launch grids are sized for), and extents that fit the packed
// rim-cell queue.
__device__ __forceinline__ bool zone_coded(const MapArgs& a, int ni, int nj, int nk) {
  return ni > 0 && nj > 0 && nk > 0 && ni < 2048 && nj < 2048 && nk < 1024 &&
         ni <= a.zone_span && nj <= a.zone_span && nk <= a.zone_span &&
         (int64_t)ni * nj * nk <= (int64_t)a.zone_cap;
}

struct SfShared {
  int qnode[3][kSfChunk];       // rim cells per column axis: li | lj << 11 | lk << 22
  int qlen[3];
  double sq[kSfBatch][3][kSfSlices];  // lane, loop, column squares per cell
  double cmm[kSfBatch][2];      // column min / max
  double off[kSfBatch][3];      // dx, dy, dz of each cell
  unsigned short upair[kSfUndecCap];  // (cell << 8) | (lane slice << 4) | loop slice
  int ulen;
  int cnt[kSfBatch];
};

__device__ __forceinline__ void sf_cell_offsets(const Geom& g, const int* z, int li, int lj,
                                                int lk, double xL, double yL, double zL,
                                                double& dx, double& dy, double& dz) {
  dx = (double)(z[0] + li + g.off[0]) - xL;
  dy = (double)(z[2] + lj + g.off[1]) - yL;
  dz = (double)(z[4] + lk + g.off[2]) - zL;
}

template <int AXIS>
__device__ __forceinline__ void sf_count_queue(const MapArgs& a, SfShared& sh, const int* z,
                                               int nj, int nk, double xL, double yL, double zL,
                                               double rSq, unsigned short* out) {
  const int tid = threadIdx.x;
  const int nq = sh.qlen[AXIS];
  const bool two_d = a.g.dim == 2;
  for (int b0 = 0; b0 < nq; b0 += kSfBatch) {
    const int nb = min(kSfBatch, nq - b0);
    // offsets of each cell (lane, loop, column order), then their squares
    if (tid < nb) {
      const int p = sh.qnode[AXIS][b0 + tid];
      double dx, dy, dz, lane, loop, col;
      sf_cell_offsets(a.g, z, p & 2047, (p >> 11) & 2047, p >> 22, xL, yL, zL, dx, dy, dz);
      sf_axis_offsets(AXIS, dx, dy, dz, lane, loop, col);
      sh.off[tid][0] = lane;
      sh.off[tid][1] = loop;
      sh.off[tid][2] = col;
    }
    if (tid < kSfBatch) sh.cnt[tid] = 0;
    if (tid == 0) sh.ulen = 0;
    __syncthreads();
    for (int t = tid; t < nb * 3 * kSfSlices; t += 256) {
      const int cell = t / (3 * kSfSlices);
      const int rem = t - cell * 3 * kSfSlices;
      const int arr = rem / kSfSlices, sl = rem - arr * kSfSlices;
      const double o = sh.off[cell][arr] + sf_delta(sl);
      sh.sq[cell][arr][sl] = o * o;
    }
    __syncthreads();
    if (tid < nb) {
      double lo = sh.sq[tid][2][0], hi = lo;
      for (int c = 1; c < kSfSlices; ++c) {
        const double v = sh.sq[tid][2][c];
        lo = v < lo ? v : lo;
        hi = v > hi ? v : hi;
      }
      sh.cmm[tid][0] = lo;
      sh.cmm[tid][1] = hi;
    }
    __syncthreads();
    if (tid < nb * kSfSlices) {
      const int cell = tid / kSfSlices, l = tid - cell * kSfSlices;
      int n;
      if (two_d) {
        n = sf_lane_count_2d(sh.sq[cell][0][l], sh.sq[cell][1], rSq);
      } else {
        unsigned undec = 0u;
        n = sf_lane_count<AXIS>(sh.sq[cell][0][l], sh.sq[cell][1], sh.cmm[cell][0],
                                sh.cmm[cell][1], rSq, undec);
        if (undec) {
          const int k0 = atomicAdd(&sh.ulen, __popc(undec));
          int k = k0;
          for (int m = 0; m < kSfSlices; ++m)
            if (undec & (1u << m)) sh.upair[k++] = (unsigned short)((cell << 8) | (l << 4) | m);
        }
      }
      atomicAdd(&sh.cnt[cell], n);
    }
    __syncthreads();
    const int nu = sh.ulen;
    for (int u = tid; u < nu; u += 256) {
      const unsigned short p = sh.upair[u];
      const int cell = p >> 8, l = (p >> 4) & 15, m = p & 15;
      atomicAdd(&sh.cnt[cell],
                sf_column_count<AXIS>(sh.sq[cell][0][l], sh.sq[cell][1][m], sh.sq[cell][2], rSq));
    }
    __syncthreads();
    if (tid < nb) {
      const int p = sh.qnode[AXIS][b0 + tid];
      out[((p & 2047) * nj + ((p >> 11) & 2047)) * nk + (p >> 22)] = (unsigned short)sh.cnt[tid];
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void k_entry_sf(const MapArgs a) {
  __shared__ SfShared sh;
  const int chunks = (a.zone_cap + kSfChunk - 1) / kSfChunk;
  const int e = blockIdx.x / chunks;
  const int base = (blockIdx.x - e * chunks) * kSfChunk;
  if (e >= a.n) return;
  const Geom& g = a.g;
  const int* z = a.zone + 6 * e;
  const int ni = z[1] - z[0], nj = z[3] - z[2], nk = z[5] - z[4];
  if (ni <= 0 || nj <= 0 || nk <= 0) return;
  if (!zone_coded(a, ni, nj, nk)) return;  // (computed inline by k_map)
  if (g.dim == 3 && nj <= kLineMax && nk <= kLineMax) return;  // (k_entry_sf_lines)
  const int tot = ni * nj * nk, njk = nj * nk;
  if (base >= tot) return;
  const double xL = a.xl[4 * e + 0], yL = a.xl[4 * e + 1], zL = a.xl[4 * e + 2];
  const double rL = a.xl[4 * e + 3];
  unsigned short* out = a.sfcode + (size_t)e * a.zone_cap;
  const int tid = threadIdx.x;
  if (tid < 3) sh.qlen[tid] = 0;
  __syncthreads();
  const int cnt = min(kSfChunk, tot - base);
  for (int t = tid; t < cnt; t += 256) {
    const int idx = base + t;
    const int li = idx / njk;
    const int rem = idx - li * njk;
    const int lj = rem / nk;
    const int lk = rem - lj * nk;
    double dx, dy, dz;
    sf_cell_offsets(g, z, li, lj, lk, xL, yL, zL, dx, dy, dz);
    const int cls = sf_classify(a.hsq_half, dx, dy, dz, rL);
    if (cls == SF_ZERO) {
      out[base + t] = 0;
    } else if (cls == SF_ONE) {
      out[base + t] = kSfCodeOne;
    } else {
      const int axis = g.dim == 2 ? 2 : sf_column_axis(dx, dy, dz);
      sh.qnode[axis][atomicAdd(&sh.qlen[axis], 1)] = li | (lj << 11) | (lk << 22);
    }
  }
  __syncthreads();
  const double rSq = rL * rL;
  sf_count_queue<2>(a, sh, z, nj, nk, xL, yL, zL, rSq, out);
  if (g.dim == 3) {
    sf_count_queue<0>(a, sh, z, nj, nk, xL, yL, zL, rSq, out);
    sf_count_queue<1>(a, sh, z, nj, nk, xL, yL, zL, rSq, out);
  }
}

// 3D solid fractions by lines (sf_first_pass / sf_last_pass): one workgroup
// per (entry, x plane of its zone).  The plane's y and z sub-point squares
// and the interval of every (a, j, b) line go through LDS; each cell of the
// plane then gets its code from the reference's early exits or, when it
// needs counting, from the summed overlaps.  Zones wider than kLineMax in y
// or z are left to k_entry_sf.
struct SfLineShared {
  double X[kSfSlices];
  double Y[kLineMax * kSfSlices];
  double Z[kLineMax * kSfSlices];
  int part[kLineMax][kLineMax + 1];
  int diff[kLineMax][kLineMax + 1];
  int pstar;
  int jlo, jhi;  // rows whose lines can reach the sphere
  double t0;   // z offset of the line's first point
};

__global__ __launch_bounds__(256) void k_entry_sf_lines(const MapArgs a, int planes) {
  __shared__ SfLineShared sh;
  // grid (planes, entries), or one dimension of entries x planes when the
  // entries exceed the grid's second dimension
  int e, li;
  if (gridDim.y > 1) {
    li = (int)blockIdx.x;
    e = (int)blockIdx.y;
  } else {
    e = (int)blockIdx.x / planes;
    li = (int)blockIdx.x - e * planes;
  }
  if (e >= a.n) return;
  const Geom& g = a.g;
  const int* z = a.zone + 6 * e;
  const int ni = z[1] - z[0], nj = z[3] - z[2], nk = z[5] - z[4];
  if (!zone_coded(a, ni, nj, nk) || li >= ni) return;
  const double xL = a.xl[4 * e + 0], yL = a.xl[4 * e + 1], zL = a.xl[4 * e + 2];
  const double rL = a.xl[4 * e + 3];
  const double rSq = rL * rL;
  unsigned short* out = a.sfcode + (size_t)e * a.zone_cap + (size_t)li * nj * nk;
  const int tid = threadIdx.x;
  const double dx = (double)(z[0] + li + g.off[0]) - xL;

  if (nj > kLineMax || nk > kLineMax) return;  // (k_entry_sf counts these)

  const int P = kSfSlices * nk;
  for (int t = tid; t < kSfSlices * (1 + nj + nk); t += 256) {
    const int sl = t % kSfSlices, row = t / kSfSlices;
    double d;
    double* dst;
    if (row == 0) {
      d = dx;
      dst = sh.X + sl;
    } else if (row <= nj) {
      d = (double)(z[2] + (row - 1) + g.off[1]) - yL;
      dst = sh.Y + kSfSlices * (row - 1) + sl;
    } else {
      d = (double)(z[4] + (row - 1 - nj) + g.off[2]) - zL;
      dst = sh.Z + kSfSlices * (row - 1 - nj) + sl;
    }
    const double o = d + sf_delta(sl);
    *dst = o * o;
    if (row == 1 + nj && sl == 0) sh.t0 = o;
  }
  for (int t = tid; t < nj * (kLineMax + 1); t += 256) {
    (&sh.part[0][0])[t] = 0;
    (&sh.diff[0][0])[t] = 0;
  }
  if (tid == 0) {
    sh.jlo = nj;
    sh.jhi = -1;
  }
  __syncthreads();
  // the line's minimum: along the V, the first point of the lowest value is
  // the only one below its predecessor and not above its successor
  for (int p = tid; p < P; p += 256) {
    const double zp = sh.Z[p];
    if ((p == 0 || sh.Z[p - 1] > zp) && (p == P - 1 || sh.Z[p + 1] >= zp)) sh.pstar = p;
  }
  __syncthreads();
  const int pstar = sh.pstar;
  const double t0 = sh.t0;
  // Rows none of whose lines can reach the sphere are skipped: a line of row
  // j has s = X[a] + Y[j, b] >= minX + minY_j, and rounding is monotone, so
  // when minX + minY_j + Z[pstar] >= r^2 every line of the row fails the
  // first test of sf_line_interval.  Lines are enumerated over the rows
  // [jlo, jhi] between the first and the last row that passes.
  if (tid < nj) {
    double mx = sh.X[0], my = sh.Y[kSfSlices * tid];
    for (int c = 1; c < kSfSlices; ++c) {
      mx = fmin(mx, sh.X[c]);
      my = fmin(my, sh.Y[kSfSlices * tid + c]);
    }
    if (mx + my + sh.Z[pstar] < rSq) {
      atomicMin(&sh.jlo, tid);
      atomicMax(&sh.jhi, tid);
    }
  }
  __syncthreads();
  const int jlo = sh.jlo, nr = sh.jhi - sh.jlo + 1;  // nr <= 0: no line reaches
  // column fastest across lanes: lanes of a wave add into different columns
  int ab = nr > 0 ? tid / nr : kSfPairs, jr = nr > 0 ? tid - ab * nr : 0;
  const int dab = nr > 0 ? 256 / nr : 0, dj = 256 - dab * nr;
  for (; ab < kSfPairs; ab += dab, jr += dj) {
    if (jr >= nr) {
      jr -= nr;
      ++ab;
      if (ab >= kSfPairs) break;
    }
    const int j = jlo + jr;
    const int ia = ab / kSfSlices, ib = ab - ia * kSfSlices;
    const double sxy = sh.X[ia] + sh.Y[kSfSlices * j + ib];
    int lo, hi;
    if (!sf_line_interval(sh.Z, P, pstar, t0, sxy, rSq, lo, hi)) continue;
    int klo, khi, nlo, nhi;
    sf_line_split(lo, hi, klo, khi, nlo, nhi);
    atomicAdd(&sh.part[j][klo], nlo);
    if (khi > klo) {
      atomicAdd(&sh.part[j][khi], nhi);
      if (khi > klo + 1) {
        atomicAdd(&sh.diff[j][klo + 1], kSfSlices);
        atomicAdd(&sh.diff[j][khi], -kSfSlices);
      }
    }
  }
  __syncthreads();
  if (tid < nj) {
    int run = 0;
    for (int k = 0; k < nk; ++k) {
      run += sh.diff[tid][k];
      sh.part[tid][k] += run;
    }
  }
  __syncthreads();
  for (int t = tid; t < nj * nk; t += 256) {
    const int j = t / nk, k = t - (t / nk) * nk;
    const double dy = (double)(z[2] + j + g.off[1]) - yL;
    const double dz = (double)(z[4] + k + g.off[2]) - zL;
    const int cls = sf_classify(a.hsq_half, dx, dy, dz, rL);
    out[t] = cls == SF_ZERO ? 0 : cls == SF_ONE ? kSfCodeOne : (unsigned short)sh.part[j][k];
  }
}

// One workgroup (256 threads = one brick of 256 nodes) per brick.  The
// entries to replay are taken in list order: the brick's own list, staged
// whole and ranked, or - when more entries touch the brick than its list
// holds - the whole particle list, staged kBrickCap at a time.  Every node
// replays them with the solid fractions from k_entry_sf.  The loads that do
// not depend on each other are issued together: count, list and the node's
// slots first, then all listed entries, then the codes of a batch of
// entries.
__device__ __forceinline__ double map_code_sf(int dim, unsigned short cd) {
  return cd == kSfCodeOne ? 1.0 : sf_from_count(dim, cd);
}

// Entries whose solid-fraction codes a node loads together before applying
// them in order (their loads are independent, so they overlap).
#ifndef LBMPSM_MAP_BATCH
#define LBMPSM_MAP_BATCH 2
#endif
constexpr int kMapBatch = LBMPSM_MAP_BATCH;
#ifndef LBMPSM_MAP_WAVES
#define LBMPSM_MAP_WAVES 8  // > 0: ask for this many waves per SIMD for k_map
#endif
#if LBMPSM_MAP_WAVES > 0
#define LBMPSM_MAP_OCC __attribute__((amdgpu_waves_per_eu(LBMPSM_MAP_WAVES)))
#else
#define LBMPSM_MAP_OCC
#endif

template <bool GENERAL, int DIM>
__device__ __forceinline__ void map_brick(const MapArgs& a, int64_t b, int bx, int by, int bz,
                                          MapEntry* ent, int* raw, int* flag) {
  // the brick's shape is fixed per dimension, so a thread's node within it
  // comes from shifts and masks
  constexpr int BI = DIM == 3 ? kBrick3[0] : kBrick2[0];
  constexpr int BJ = DIM == 3 ? kBrick3[1] : kBrick2[1];
  constexpr int BK = DIM == 3 ? kBrick3[2] : kBrick2[2];
  const int tid = threadIdx.x;
  // independent loads first: the brick's count and list, the node's slots
  const int m = a.count[b];
  const int rv = tid < kBrickCap ? a.list[b * kBrickCap + tid] : 0;
  if (m == 0) return;
  if (!GENERAL && m > kBrickCap) {  // longer than its list: the general kernel
    if (tid == 0) a.slow_list[atomicAdd(a.slow_count, 1)] = (int)b;
    return;
  }
  const Geom& g = a.g;
  const int64_t plane = g.plane;
  const bool listed = m <= kBrickCap;

  const int lk = tid % BK;
  const int lj = (tid / BK) % BJ;
  const int li = tid / (BK * BJ);
  const int i = bx * BI + li, j = by * BJ + lj, k = bz * BK + lk;
  const bool valid = i < g.sx && j < g.sy && k < g.sz;
  const int n = valid ? (i * g.sy + j) * g.sz + k : 0;
  NodeSlots S;
  S.t0 = valid ? a.s.tag[n] : 0;
  S.t1 = valid ? a.s.tag[plane + n] : 0;
  // the fresh bits of an earlier map whose step has not run yet stay with
  // the slots this map leaves alone
  const uint8_t m_old = valid ? a.s.mask[n] : (uint8_t)0;
  S.d0 = S.d1 = false;
  S.sf0 = S.sf1 = 0.0;
  // reference (halo'd) node indices, as used in the geometry
  const int ri = i + g.off[0], rj = j + g.off[1], rk = k + g.off[2];

  if (listed) {
    // stage every listed entry at once, in list order: the lanes load their
    // entries, rank the list (entries are distinct: rank = number of smaller
    // entries) and store each entry at its rank
    MapEntry L;
    if (tid < m) {
      raw[tid] = rv;
      for (int c = 0; c < 6; ++c) L.zone[c] = a.zone[6 * rv + c];
      L.tag = a.tag[rv];
      L.raw = rv;
      for (int c = 0; c < 4; ++c) L.xl[c] = a.xl[4 * rv + c];
      for (int c = 0; c < 6; ++c) L.kin[c] = a.kin[6 * rv + c];
    }
    if (!GENERAL && tid == 0) *flag = 0;
    __syncthreads();
    if (tid < m) {
      int r = 0;
      for (int u = 0; u < m; ++u) r += raw[u] < rv;
      const int* z = L.zone;
      L.coded = zone_coded(a, z[1] - z[0], z[3] - z[2], z[5] - z[4]) ? 1 : 0;
      if (!GENERAL && !L.coded) *flag = 1;
      ent[r] = L;
    }
    __syncthreads();
    // an entry without codes: the general kernel takes the whole brick
    if (!GENERAL && *flag) {
      if (tid == 0) a.slow_list[atomicAdd(a.slow_count, 1)] = (int)b;
      return;
    }
    if (valid) {
      for (int r0 = 0; r0 < m; r0 += kMapBatch) {
        unsigned short cd[kMapBatch];
        unsigned in = 0u, coded = 0u;
#pragma unroll
        for (int u = 0; u < kMapBatch; ++u) {
          cd[u] = 0;
          if (r0 + u < m) {
            const MapEntry& E = ent[r0 + u];
            const int z0 = E.zone[0], z1 = E.zone[1], z2 = E.zone[2], z3 = E.zone[3];
            const int z4 = E.zone[4], z5 = E.zone[5];
            if (in_range(i, z0, z1) & in_range(j, z2, z3) & in_range(k, z4, z5)) {
              in |= 1u << u;
              if (E.coded) {
                coded |= 1u << u;
                // the reference's early exits decide most cells (and give
                // the codes the solid-fraction kernel stored for them); only
                // a cell that needs its count loads the code
                const int cls = sf_classify(a.hsq_half, (double)ri - E.xl[0],
                                            (double)rj - E.xl[1], (double)rk - E.xl[2], E.xl[3]);
                if (cls == SF_ONE) {
                  cd[u] = kSfCodeOne;
                } else if (cls == SF_COUNT) {
                  const int nj = z3 - z2, nk = z5 - z4;
                  const int idx = ((i - z0) * nj + (j - z2)) * nk + (k - z4);
                  cd[u] = a.sfcode[(size_t)E.raw * a.zone_cap + idx];
                }
              }
            }
          }
        }
#pragma unroll
        for (int u = 0; u < kMapBatch; ++u) {
          if (!(in & (1u << u))) continue;
          const MapEntry& E = ent[r0 + u];
          double sf;
          if (!GENERAL || (coded & (1u << u)))
            sf = map_code_sf(g.dim, cd[u]);
          else
            sf = solid_fraction(g.dim, a.hsq_half, ri, rj, rk, E.xl[0], E.xl[1], E.xl[2], E.xl[3]);
          map_transition(a, E, sf, ri, rj, rk, n, S);
        }
      }
    }
    __syncthreads();
  } else if constexpr (GENERAL) {
    // more entries touch the brick than its list holds: replay the whole
    // particle list, staged kBrickCap entries at a time
    for (int base = 0; base < a.n; base += kBrickCap) {
      const int cnt = min(kBrickCap, a.n - base);
      if (tid < cnt) {
        const int v = base + tid;
        MapEntry& E = ent[tid];
        for (int c = 0; c < 6; ++c) E.zone[c] = a.zone[6 * v + c];
        E.tag = a.tag[v];
        for (int c = 0; c < 4; ++c) E.xl[c] = a.xl[4 * v + c];
        for (int c = 0; c < 6; ++c) E.kin[c] = a.kin[6 * v + c];
      }
      __syncthreads();
      if (valid) {
        for (int t = 0; t < cnt; ++t) {
          const MapEntry& E = ent[t];
          const int* z = E.zone;
          if (i < z[0] || i >= z[1] || j < z[2] || j >= z[3] || k < z[4] || k >= z[5]) continue;
          const int nj = z[3] - z[2], nk = z[5] - z[4];
          double sf;
          if (zone_coded(a, z[1] - z[0], nj, nk)) {
            const int idx = ((i - z[0]) * nj + (j - z[2])) * nk + (k - z[4]);
            sf = map_code_sf(g.dim, a.sfcode[(size_t)(base + t) * a.zone_cap + idx]);
          } else {
            sf = solid_fraction(g.dim, a.hsq_half, ri, rj, rk, E.xl[0], E.xl[1], E.xl[2],
                                E.xl[3]);
          }
          map_transition(a, E, sf, ri, rj, rk, n, S);
        }
      }
      __syncthreads();
    }
  }
  // the brick list is consumed: leave its count at zero for the next map
  if (tid == 0) a.count_reset[b] = 0;
  if (!valid) return;
  // a slot cleared this step (setToZero) keeps no force
  if (S.d0) {
    a.s.tag[n] = S.t0;
    if (a.s.sff) a.s.sff[n] = (float)S.sf0;
    else a.s.sf[n] = S.sf0;
    if (S.t0 <= 0) store_node_vec(a.s.Fh, a.s.Fhf, n, plane, 0.0, 0.0, 0.0);
  }
  if (S.d1) {
    a.s.tag[plane + n] = S.t1;
    if (a.s.sff) a.s.sff[plane + n] = (float)S.sf1;
    else a.s.sf[plane + n] = S.sf1;
    if (S.t1 <= 0) store_node_vec(a.s.Fh, a.s.Fhf, 3 * plane + n, plane, 0.0, 0.0, 0.0);
  }
  // a slot set this step has its force reset (src/LBM_PSM_lattice.cpp:
  // 257-299): mark it fresh (mask bits 2, 3) and the collide starts its
  // force from zero instead of reading it.  A slot this map did not write
  // keeps its fresh bit: it was set by an earlier map since the last step
  // (the caller set two lists in a row), so its force is still zero.
  if (S.d0 || S.d1) {
    const uint8_t f0 = S.d0 ? (S.t0 > 0 ? kMaskFresh0 : 0) : (m_old & kMaskFresh0);
    const uint8_t f1 = S.d1 ? (S.t1 > 0 ? kMaskFresh1 : 0) : (m_old & kMaskFresh1);
    a.s.mask[n] = (uint8_t)((S.t0 > 0 ? kMaskOcc0 : 0) | (S.t1 > 0 ? kMaskOcc1 : 0) | f0 | f1);
  }
}

// The common case: bricks whose entries fit their list and all have codes.
// Other bricks go, untouched, to a work list for k_map_general.
// GRID3: the grid is (z, y, x) bricks; else one dimension of bricks.
template <int DIM, bool GRID3>
__global__ __launch_bounds__(256) LBMPSM_MAP_OCC void k_map(const MapArgs a) {
  __shared__ MapEntry ent[kBrickCap];  // in list order
  __shared__ int raw[kBrickCap];
  __shared__ int flag;
  int bx, by, bz;
  int64_t b;
  if constexpr (GRID3) {
    bz = blockIdx.x;
    by = blockIdx.y;
    bx = blockIdx.z;
    b = ((int64_t)bx * a.nbj + by) * a.nbk + bz;
  } else {
    b = blockIdx.x;
    bx = (int)(b / ((int64_t)a.nbj * a.nbk));
    by = (int)((b / a.nbk) % a.nbj);
    bz = (int)(b % a.nbk);
  }
  map_brick<false, DIM>(a, b, bx, by, bz, ent, raw, &flag);
}

// The bricks k_map passed on: overflowing lists and entries whose zones
// have no codes (solid fractions computed here, cell by cell).  A fixed grid
// walks the work list; an empty list costs one load per workgroup.
constexpr int kMapGeneralBlocks = 256;
template <int DIM>
__global__ __launch_bounds__(256) void k_map_general(const MapArgs a) {
  __shared__ MapEntry ent[kBrickCap];
  __shared__ int raw[kBrickCap];
  const int cnt = *a.slow_count;
  for (int w = blockIdx.x; w < cnt; w += gridDim.x) {
    const int64_t b = a.slow_list[w];
    const int bx = (int)(b / ((int64_t)a.nbj * a.nbk));
    const int by = (int)((b / a.nbk) % a.nbj);
    const int bz = (int)(b % a.nbk);
    map_brick<true, DIM>(a, b, bx, by, bz, ent, raw, nullptr);
    __syncthreads();
  }
}

// Force, torque and stresslet of one list entry, split over kHydroGroups
// workgroups of kHydroThreads threads.  Group g takes the x-planes
// [g ni / G, (g+1) ni / G) of the entry's reduction zone; a thread takes (y, z)
// columns of the zone and, for each, loads the slot tags of all its planes
// together (kHydroU at a time), then the node forces of the matching slots.
// The sums are combined in a fixed order: a shuffle tree within each wave,
// the wave sums in wave order, then (k_hydro_sum) the group sums in group
// order.  Terms as src/LBM_PSM_exchangeParticleData.cpp:197-240:
// F += Fhyd ff, T += (r x Fhyd) tf, S -= 1/2 (r_a F_b + r_b F_a) tf with r
// the node offset from the entry's lattice position.
#ifndef LBMPSM_HYDRO_THREADS
#define LBMPSM_HYDRO_THREADS 256
#endif
#ifndef LBMPSM_HYDRO_U
#define LBMPSM_HYDRO_U 8
#endif
constexpr int kHydroThreads = LBMPSM_HYDRO_THREADS;
constexpr int kHydroU = LBMPSM_HYDRO_U;
static_assert(kHydroThreads % 64 == 0 && kHydroThreads <= 1024, "whole waves");

// FT: type of the stored node forces (double for FP64 storage, else float)
template <typename FT>
__global__ __launch_bounds__(kHydroThreads) void k_hydro(const HydroArgs a, const FT* Fh) {
  const int p = (int)blockIdx.x / kHydroGroups;
  const int grp = (int)blockIdx.x % kHydroGroups;
  if (p >= a.n) return;
  const Geom& g = a.g;
  const int64_t plane = g.plane;
  const int* z = a.rzone + 6 * p;
  const int tag = a.tag[p];
  const double xL = a.xl[4 * p + 0], yL = a.xl[4 * p + 1], zL = a.xl[4 * p + 2];
  const double ff = a.forceFactor, tf = a.torqueFactor;
  double acc[12];
  for (int c = 0; c < 12; ++c) acc[c] = 0.0;
  const int ni = z[1] - z[0], nj = z[3] - z[2], nk = z[5] - z[4];
  const bool empty = ni <= 0 || nj <= 0 || nk <= 0;
  const int plo = empty ? 0 : (grp * ni) / kHydroGroups;
  const int phi = empty ? 0 : ((grp + 1) * ni) / kHydroGroups;
  const int npairs = empty ? 0 : nj * nk;
  const int psz = g.sy * g.sz;  // nodes per x-plane (32-bit: fewer than 2^31 stored nodes)
  constexpr int U = kHydroU;
  for (int pr = (int)threadIdx.x; pr < npairs; pr += kHydroThreads) {
    const int lj = pr / nk, lk = pr - lj * nk;
    const int j = z[2] + lj, k = z[4] + lk;
    const int nb = j * g.sz + k;
    const double dy = (double)(j + g.off[1]) - yL;
    const double dz = (double)(k + g.off[2]) - zL;
    for (int i0 = plo; i0 < phi; i0 += U) {
      int sl[U], nn[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        sl[u] = -1;
        nn[u] = 0;
        if (i0 + u < phi) nn[u] = (z[0] + i0 + u) * psz + nb;
      }
      // the occupancy bits first: a slot without its bit holds tag 0, so
      // only the tags of occupied slots are loaded (most nodes hold one
      // particle at most, and the zone's corners none)
      uint8_t mk[U];
#pragma unroll
      for (int u = 0; u < U; ++u) mk[u] = i0 + u < phi ? a.mask[nn[u]] : (uint8_t)0;
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (i0 + u < phi) {
          const int a0 = (mk[u] & kMaskOcc0) ? a.slot_tag[nn[u]] : 0;
          const int a1 = (mk[u] & kMaskOcc1) ? a.slot_tag[plane + nn[u]] : 0;
          sl[u] = tag == a0 ? 0 : (tag == a1 ? 1 : -1);
        }
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (sl[u] < 0) continue;
        const int s = sl[u], n = nn[u];
        const FT* Fs = Fh + (3 * s) * plane + n;
        const double F0 = Fs[0], F1 = Fs[plane], F2 = Fs[2 * plane];
        acc[0] += F0 * ff;
        acc[1] += F1 * ff;
        acc[2] += F2 * ff;
        const double dx = (double)(z[0] + i0 + u + g.off[0]) - xL;
        acc[3] += (dy * F2 - dz * F1) * tf;
        acc[4] += (dz * F0 - dx * F2) * tf;
        acc[5] += (dx * F1 - dy * F0) * tf;
        acc[6] -= 0.5 * (dx * F0 + dx * F0) * tf;
        acc[7] -= 0.5 * (dy * F1 + dy * F1) * tf;
        acc[8] -= 0.5 * (dz * F2 + dz * F2) * tf;
        acc[9] -= 0.5 * (dy * F0 + dx * F1) * tf;
        acc[10] -= 0.5 * (dz * F0 + dx * F2) * tf;
        acc[11] -= 0.5 * (dz * F1 + dy * F2) * tf;
      }
    }
  }
#pragma unroll
  for (int c = 0; c < 12; ++c)
    for (int off = 32; off > 0; off >>= 1) acc[c] += __shfl_xor(acc[c], off, 64);
  __shared__ double wsum[kHydroThreads / 64][12];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0)
    for (int c = 0; c < 12; ++c) wsum[wave][c] = acc[c];
  __syncthreads();
  if (threadIdx.x != 0) return;
  double gs[12];
  for (int c = 0; c < 12; ++c) {
    double s = 0.0;
    for (int w = 0; w < kHydroThreads / 64; ++w) s += wsum[w][c];
    gs[c] = s;
  }
  double* mine = a.part + ((size_t)p * kHydroGroups + grp) * 12;
  for (int c = 0; c < 12; ++c) mine[c] = gs[c];
}

// Per entry and component: the group sums in group order.
__global__ __launch_bounds__(256) void k_hydro_sum(const HydroArgs a) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= 12 * a.n) return;
  const int p = t / 12, c = t - 12 * p;
  const double* all = a.part + (size_t)p * kHydroGroups * 12;
  double s = 0.0;
  for (int q = 0; q < kHydroGroups; ++q) s += all[q * 12 + c];
  a.out[t] = s;
}

__device__ __forceinline__ bool interior_of(const Geom& g, int64_t n, int64_t& o) {
  const int k = (int)(n % g.sz);
  const int j = (int)((n / g.sz) % g.sy);
  const int i = (int)(n / ((int64_t)g.sz * g.sy));
  const int gi = i + g.off[0] - 1;
  const int gj = j + g.off[1] - 1;
  const int gk = (g.dim == 3) ? k + g.off[2] - 1 : 0;
  if (gi < 0 || gi >= g.nx || gj < 0 || gj >= g.ny || gk < 0 || gk >= g.nz) return false;
  o = ((int64_t)gi * g.ny + gj) * g.nz + gk;
  return true;
}

__global__ __launch_bounds__(256) void k_slots_export(const Geom g, const SlotsDev s, int* tag,
                                                      double* sf, double* us, double* F) {
  const int64_t n = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (n >= g.nodes) return;
  int64_t o;
  if (!interior_of(g, n, o)) return;
  const int64_t P = g.plane;
  // a slot set since the last collide reads zero force, as the reference's
  // set resets it (src/LBM_PSM_lattice.cpp:257-299); the collide starts such
  // a force from zero without the map storing one
  const uint8_t m = s.mask[n];
  for (int c = 0; c < 2; ++c) {
    const bool fresh = m & (c ? kMaskFresh1 : kMaskFresh0);
    if (tag) tag[2 * o + c] = s.tag[c * P + n];
    if (sf) sf[2 * o + c] = s.sff ? (double)s.sff[c * P + n] : s.sf[c * P + n];
    for (int d = 0; d < 3; ++d) {
      const int64_t v = (3 * c + d) * P + n;
      if (us) us[6 * o + 3 * c + d] = s.usf ? (double)s.usf[v] : s.us[v];
      if (F) F[6 * o + 3 * c + d] = fresh ? 0.0 : s.Fhf ? (double)s.Fhf[v] : s.Fh[v];
    }
  }
}

__global__ __launch_bounds__(256) void k_macros_export(const Geom g, const double* rho,
                                                       const double* u, double* ro, double* uo) {
  const int64_t n = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (n >= g.nodes) return;
  int64_t o;
  if (!interior_of(g, n, o)) return;
  if (ro) ro[o] = rho[n];
  if (uo)
    for (int c = 0; c < 3; ++c) uo[3 * o + c] = u[c * g.plane + n];
}

__global__ __launch_bounds__(256) void k_slots_reset(const Geom g, const SlotsDev s) {
  const int64_t n = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (n >= g.plane) return;
  const int64_t P = g.plane;
  for (int c = 0; c < 2; ++c) {
    s.tag[c * P + n] = 0;
    s.sf[c * P + n] = 0.0;
  }
  for (int c = 0; c < 6; ++c) {
    s.us[c * P + n] = 0.0;
    s.Fh[c * P + n] = 0.0;
  }
  s.mask[n] = 0;
}

// ---------------------------------------------------------------------------
static inline dim3 grid_for(int64_t n, int b) { return dim3((unsigned)((n + b - 1) / b)); }

hipError_t launch_fill_int(int* p, int64_t n, int value, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_int, grid_for(n, 256), dim3(256), 0, s, p, n, value);
  return hipGetLastError();
}

hipError_t launch_prep_bin(const PrepArgs& a, hipStream_t s) {
  if (a.n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_prep_bin, dim3(a.n), dim3(64), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_entry_sf(const MapArgs& a, hipStream_t s) {
  if (a.n <= 0 || a.zone_cap <= 0) return hipSuccess;
  if (a.g.dim == 3) {
    // every coded zone is at most zone_span planes wide in x
    const int planes = std::max(1, std::min(a.zone_span, a.g.sx));
    const dim3 grid = a.n > 1 && a.n <= 65535 ? dim3((unsigned)planes, (unsigned)a.n)
                                              : dim3((unsigned)((int64_t)a.n * planes));
    hipLaunchKernelGGL(k_entry_sf_lines, grid, dim3(256), 0, s, a, planes);
  }
  if (a.g.dim == 2 || a.zone_span > kLineMax) {  // zones the line kernel leaves
    const int64_t chunks = (a.zone_cap + kSfChunk - 1) / kSfChunk;
    hipLaunchKernelGGL(k_entry_sf, dim3((unsigned)(a.n * chunks)), dim3(256), 0, s, a);
  }
  return hipGetLastError();
}

hipError_t launch_map(const MapArgs& a, int64_t nbricks, hipStream_t s) {
  // without entries no node changes state (and k_prep_bin, which zeroes the
  // general kernel's work list, did not run)
  if (nbricks <= 0 || a.n <= 0) return hipSuccess;
  // the brick shapes the kernels are compiled for
  const bool d3 = a.g.dim == 3;
  const int* br = d3 ? kBrick3 : kBrick2;
  if (a.bi != br[0] || a.bj != br[1] || a.bk != br[2]) return hipErrorInvalidValue;
  const bool grid3 = a.nbj <= 65535 && a.nbi <= 65535;
  const dim3 g3((unsigned)a.nbk, (unsigned)a.nbj, (unsigned)a.nbi), g1((unsigned)nbricks);
  const dim3 gg((unsigned)std::min<int64_t>(nbricks, kMapGeneralBlocks));
  if (d3) {
    if (grid3) hipLaunchKernelGGL((k_map<3, true>), g3, dim3(256), 0, s, a);
    else hipLaunchKernelGGL((k_map<3, false>), g1, dim3(256), 0, s, a);
    hipLaunchKernelGGL(k_map_general<3>, gg, dim3(256), 0, s, a);
  } else {
    if (grid3) hipLaunchKernelGGL((k_map<2, true>), g3, dim3(256), 0, s, a);
    else hipLaunchKernelGGL((k_map<2, false>), g1, dim3(256), 0, s, a);
    hipLaunchKernelGGL(k_map_general<2>, gg, dim3(256), 0, s, a);
  }
  return hipGetLastError();
}

hipError_t launch_hydro(const HydroArgs& a, hipStream_t s) {
  if (a.n <= 0) return hipSuccess;
  const dim3 grid((unsigned)a.n * kHydroGroups);
  if (a.Fhf)
    hipLaunchKernelGGL(k_hydro<float>, grid, dim3(kHydroThreads), 0, s, a, a.Fhf);
  else
    hipLaunchKernelGGL(k_hydro<double>, grid, dim3(kHydroThreads), 0, s, a, a.Fh);
  hipLaunchKernelGGL(k_hydro_sum, grid_for(12 * (int64_t)a.n, 256), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_slots_export(const Geom& g, const SlotsDev& sl, int* tag, double* sf,
                               double* us, double* F, hipStream_t s) {
  hipLaunchKernelGGL(k_slots_export, grid_for(g.nodes, 256), dim3(256), 0, s, g, sl, tag, sf, us, F);
  return hipGetLastError();
}

hipError_t launch_macros_export(const Geom& g, const double* rho, const double* u,
                                double* rho_out, double* u_out, hipStream_t s) {
  hipLaunchKernelGGL(k_macros_export, grid_for(g.nodes, 256), dim3(256), 0, s, g, rho, u, rho_out, u_out);
  return hipGetLastError();
}

hipError_t launch_slots_reset(const Geom& g, const SlotsDev& sl, hipStream_t s) {
  hipLaunchKernelGGL(k_slots_reset, grid_for(g.plane, 256), dim3(256), 0, s, g, sl);
  return hipGetLastError();
}

}  // namespace lbmpsm
