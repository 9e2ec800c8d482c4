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
  if (p >= a.n) return;
  double loc[3], r_lb;
  int z[6], rz[6];
  entry_geometry(a, p, loc, r_lb, z, rz);
  if (threadIdx.x == 0) {
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
  double xl[4];
  double kin[6];
};

// One state-machine transition of a node for one entry whose zone contains
// it, given the entry's solid fraction at the node,
// src/LBM_PSM_exchangeParticleData.cpp:69-106.  Solid fractions of the
// node's slots are read from memory only when the entry's tag is found in
// a slot (the only case that uses them).
struct NodeSlots {
  int t0, t1;
  double sf0, sf1;
  double u0[3], u1[3];
  bool d0, d1;      // slot written this step
  bool k0, k1;      // sf0 / sf1 known (loaded or written)
};

__device__ __forceinline__ void map_transition(const MapArgs& a, const MapEntry& E, double sf,
                                               int ri, int rj, int rk, int n, NodeSlots& S) {
  const int64_t plane = a.g.plane;
  const int tag = E.tag;
  int id_old = 0;
  double sf_old = 0.0;
  if (tag == S.t0) {
    if (!S.k0) {
      S.sf0 = a.s.sf[n];
      S.k0 = true;
    }
    id_old = S.t0;
    sf_old = S.sf0;
  } else if (tag == S.t1) {
    if (!S.k1) {
      S.sf1 = a.s.sf[plane + n];
      S.k1 = true;
    }
    id_old = S.t1;
    sf_old = S.sf1;
  }
  const int dec = (sf > 0.00001) + 2 * (sf_old > 0.00001);
  bool do_set = false, do_clear = false;
  switch (dec) {
    case 0: do_clear = true; break;
    case 1: do_set = true; break;
    case 2: do_clear = (id_old == tag); break;
    case 3: do_set = (sf > sf_old || id_old == tag); break;
  }
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
    if (sl == 0) {
      S.t0 = tag; S.sf0 = sf; S.k0 = true; S.d0 = true;
      S.u0[0] = un[0]; S.u0[1] = un[1]; S.u0[2] = un[2];
    } else if (sl == 1) {
      S.t1 = tag; S.sf1 = sf; S.k1 = true; S.d1 = true;
      S.u1[0] = un[0]; S.u1[1] = un[1]; S.u1[2] = un[2];
    }
  } else if (do_clear) {  // setToZero, src/LBM_PSM_lattice.cpp:302-324
    if (S.t0 == tag) {
      S.t0 = 0; S.sf0 = 0.0; S.k0 = true; S.d0 = true;
      S.u0[0] = S.u0[1] = S.u0[2] = 0.0;
    } else if (S.t1 == tag) {
      S.t1 = 0; S.sf1 = 0.0; S.k1 = true; S.d1 = true;
      S.u1[0] = S.u1[1] = S.u1[2] = 0.0;
    }
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
constexpr int kSfBatch = 24;
constexpr int kSfUndecCap = kSfBatch * kSfPairs;
static_assert(kSfBatch * kSfSlices <= 256, "one lane per cell slice");

// Zones whose codes k_entry_sf stores (others are computed inline by k_map):
// at most zone_cap nodes, and extents that fit the packed rim-cell queue.
__device__ __forceinline__ bool zone_coded(int ni, int nj, int nk, int zone_cap) {
  return ni > 0 && nj > 0 && nk > 0 && ni < 2048 && nj < 2048 && nk < 1024 &&
         (int64_t)ni * nj * nk <= (int64_t)zone_cap;
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
  if (!zone_coded(ni, nj, nk, a.zone_cap)) return;  // (computed inline by k_map)
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

// One workgroup (256 threads = one brick of 256 nodes) per brick.  The
// entries to replay are taken in list order: the brick's own list, sorted,
// or - when more entries touch the brick than its list holds - the whole
// particle list, staged kMapChunk at a time.  Every node replays them with
// the solid fractions from k_entry_sf.
constexpr int kMapChunk = 32;

__global__ __launch_bounds__(256) void k_map(const MapArgs a) {
  __shared__ int order[kBrickCap];
  __shared__ int raw[kBrickCap];
  __shared__ MapEntry ent[kMapChunk];
  __shared__ int entid[kMapChunk];
  const int64_t b = blockIdx.x;
  const int m = a.count[b];
  if (m == 0) return;
  const Geom& g = a.g;
  const int64_t plane = g.plane;
  const bool listed = m <= kBrickCap;
  const int tid = threadIdx.x;

  const int bb = (int)b;
  const int bx = bb / (a.nbj * a.nbk);
  const int by = (bb / a.nbk) % a.nbj;
  const int bz = bb % a.nbk;
  const int lk = tid % a.bk;
  const int lj = (tid / a.bk) % a.bj;
  const int li = tid / (a.bk * a.bj);
  const int i = bx * a.bi + li, j = by * a.bj + lj, k = bz * a.bk + lk;
  const bool valid = i < g.sx && j < g.sy && k < g.sz;
  const int n = valid ? (i * g.sy + j) * g.sz + k : 0;
  NodeSlots S;
  S.t0 = valid ? a.s.tag[n] : 0;
  S.t1 = valid ? a.s.tag[plane + n] : 0;
  S.sf0 = S.sf1 = 0.0;
  S.d0 = S.d1 = S.k0 = S.k1 = false;
  for (int c = 0; c < 3; ++c) S.u0[c] = S.u1[c] = 0.0;
  // reference (halo'd) node indices, as used in the geometry
  const int ri = i + g.off[0], rj = j + g.off[1], rk = k + g.off[2];

  if (listed) {
    // entries are distinct, so rank = number of smaller entries
    if (tid < m) raw[tid] = a.list[b * kBrickCap + tid];
    __syncthreads();
    if (tid < m) {
      const int v = raw[tid];
      int r = 0;
      for (int u = 0; u < m; ++u) r += raw[u] < v;
      order[r] = v;
    }
    __syncthreads();
  }
  const int total = listed ? m : a.n;
  for (int base = 0; base < total; base += kMapChunk) {
    const int cnt = min(kMapChunk, total - base);
    if (tid < cnt) {
      const int v = listed ? order[base + tid] : base + tid;
      MapEntry& E = ent[tid];
      for (int c = 0; c < 6; ++c) E.zone[c] = a.zone[6 * v + c];
      E.tag = a.tag[v];
      for (int c = 0; c < 4; ++c) E.xl[c] = a.xl[4 * v + c];
      for (int c = 0; c < 6; ++c) E.kin[c] = a.kin[6 * v + c];
      entid[tid] = v;
    }
    __syncthreads();
    if (valid) {
      for (int t = 0; t < cnt; ++t) {
        const MapEntry& E = ent[t];
        const int* z = E.zone;
        if (i < z[0] || i >= z[1] || j < z[2] || j >= z[3] || k < z[4] || k >= z[5]) continue;
        const int nj = z[3] - z[2], nk = z[5] - z[4];
        double sf;
        if (zone_coded(z[1] - z[0], nj, nk, a.zone_cap)) {
          const int idx = ((i - z[0]) * nj + (j - z[2])) * nk + (k - z[4]);
          const unsigned short cd = a.sfcode[(size_t)entid[t] * a.zone_cap + idx];
          sf = cd == kSfCodeOne ? 1.0 : sf_from_count(g.dim, cd);
        } else {
          sf = solid_fraction(g.dim, a.hsq_half, ri, rj, rk, E.xl[0], E.xl[1], E.xl[2], E.xl[3]);
        }
        map_transition(a, E, sf, ri, rj, rk, n, S);
      }
    }
    __syncthreads();
  }
  // the brick list is consumed: leave its count at zero for the next map
  if (tid == 0) a.count_reset[b] = 0;
  if (!valid) return;
  if (S.d0) {
    a.s.tag[n] = S.t0;
    a.s.sf[n] = S.sf0;
    a.s.us[n] = S.u0[0];
    a.s.us[plane + n] = S.u0[1];
    a.s.us[2 * plane + n] = S.u0[2];
    a.s.Fh[n] = 0.0;
    a.s.Fh[plane + n] = 0.0;
    a.s.Fh[2 * plane + n] = 0.0;
  }
  if (S.d1) {
    a.s.tag[plane + n] = S.t1;
    a.s.sf[plane + n] = S.sf1;
    a.s.us[3 * plane + n] = S.u1[0];
    a.s.us[4 * plane + n] = S.u1[1];
    a.s.us[5 * plane + n] = S.u1[2];
    a.s.Fh[3 * plane + n] = 0.0;
    a.s.Fh[4 * plane + n] = 0.0;
    a.s.Fh[5 * plane + n] = 0.0;
  }
  if (S.d0 || S.d1) a.s.mask[n] = (uint8_t)((S.t0 > 0 ? 1 : 0) | (S.t1 > 0 ? 2 : 0));
}

// Force, torque and stresslet of one list entry; one workgroup per entry.
__global__ __launch_bounds__(256) void k_hydro(const HydroArgs a) {
  const int p = blockIdx.x;
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
  const int tot = (ni > 0 && nj > 0 && nk > 0) ? ni * nj * nk : 0;
  const int njk = nj * nk;
  // U nodes per thread per pass so that their loads are in flight together;
  // 32-bit index arithmetic (a zone holds far fewer than 2^31 nodes)
  constexpr int U = 8;
  for (int base = threadIdx.x; base < tot; base += 256 * U) {
    int nn[U], sl[U], ii[U], jj[U], kk[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int t = base + 256 * u;
      sl[u] = -1;
      nn[u] = 0;
      ii[u] = jj[u] = kk[u] = 0;
      if (t < tot) {
        const int li = t / njk;
        const int rem = t - li * njk;
        const int lj = rem / nk;
        const int lk = rem - lj * nk;
        ii[u] = z[0] + li;
        jj[u] = z[2] + lj;
        kk[u] = z[4] + lk;
        nn[u] = (ii[u] * g.sy + jj[u]) * g.sz + kk[u];
        const int a0 = a.slot_tag[nn[u]], a1 = a.slot_tag[plane + nn[u]];
        sl[u] = tag == a0 ? 0 : (tag == a1 ? 1 : -1);
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (sl[u] < 0) continue;
      const int s = sl[u], n = nn[u];
      const double F0 = a.Fh[(3 * s) * plane + n];
      const double F1 = a.Fh[(3 * s + 1) * plane + n];
      const double F2 = a.Fh[(3 * s + 2) * plane + n];
      acc[0] += F0 * ff;
      acc[1] += F1 * ff;
      acc[2] += F2 * ff;
      const double dx = (double)(ii[u] + g.off[0]) - xL;
      const double dy = (double)(jj[u] + g.off[1]) - yL;
      const double dz = (double)(kk[u] + g.off[2]) - zL;
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
  __shared__ double red[12][256];
  for (int c = 0; c < 12; ++c) red[c][threadIdx.x] = acc[c];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w)
      for (int c = 0; c < 12; ++c) red[c][threadIdx.x] += red[c][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < 12) a.out[12 * p + threadIdx.x] = red[threadIdx.x][0];
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
  for (int c = 0; c < 2; ++c) {
    if (tag) tag[2 * o + c] = s.tag[c * P + n];
    if (sf) sf[2 * o + c] = s.sf[c * P + n];
  }
  for (int c = 0; c < 6; ++c) {
    if (us) us[6 * o + c] = s.us[c * P + n];
    if (F) F[6 * o + c] = s.Fh[c * P + n];
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

// Device-resident particle arrays into the packed per-call block
// (x[n][3], v[n][3], omega[n][3], radius[n], tag[n]).
__global__ __launch_bounds__(256) void k_pack_particles(int n, const int* tag, const double* x,
                                                        const double* v, const double* w,
                                                        const double* r, char* raw) {
  const int t = (int)blockIdx.x * 256 + (int)threadIdx.x;
  double* d = (double*)raw;
  if (t < 3 * n) {
    d[t] = x[t];
    d[3 * n + t] = v[t];
    d[6 * n + t] = w[t];
  }
  if (t < n) {
    d[9 * n + t] = r[t];
    ((int*)(raw + 80 * (size_t)n))[t] = tag[t];
  }
}

hipError_t launch_pack_particles(int n, const int* tag, const double* x, const double* v,
                                 const double* w, const double* r, char* raw, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_pack_particles, dim3((unsigned)((3 * n + 255) / 256)), dim3(256), 0, s, n,
                     tag, x, v, w, r, raw);
  return hipGetLastError();
}

hipError_t launch_entry_sf(const MapArgs& a, hipStream_t s) {
  if (a.n <= 0 || a.zone_cap <= 0) return hipSuccess;
  const int64_t chunks = (a.zone_cap + kSfChunk - 1) / kSfChunk;
  hipLaunchKernelGGL(k_entry_sf, dim3((unsigned)(a.n * chunks)), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_map(const MapArgs& a, int64_t nbricks, hipStream_t s) {
  if (nbricks <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_map, dim3((unsigned)nbricks), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_hydro(const HydroArgs& a, hipStream_t s) {
  if (a.n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_hydro, dim3(a.n), dim3(256), 0, s, a);
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
