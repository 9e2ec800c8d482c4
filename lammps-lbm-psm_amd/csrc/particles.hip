// Particle <-> lattice coupling kernels.
//
//  * prep: per list entry, the reference's lattice coordinates, radius,
//    node-velocity terms and zones (src/LBM_PSM_exchangeParticleData.cpp:31-55
//    and :175-195), plus a count of the node bricks each map zone touches.
//  * scan / fill: per-brick lists of the entries whose zone touches it.
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

__global__ __launch_bounds__(256) void k_prep_count(const PrepArgs a) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= a.n) return;
  const Geom& g = a.g;
  // src/LBM_PSM_exchangeParticleData.cpp:31-39 on one rank: the local shift
  // (nTotal - coord*2*envelope - 1) evaluates to -1.0 on every axis.
  const double off = (0 - ((double)0 * 2.0 * (double)1) - 1);
  double loc[3];
  for (int d = 0; d < 3; ++d) loc[d] = (a.x[3 * p + d] - a.origin[d]) / a.dx - off;
  const double r_lb = a.r[p] / a.dx;
  a.xl[4 * p + 0] = loc[0];
  a.xl[4 * p + 1] = loc[1];
  a.xl[4 * p + 2] = loc[2];
  a.xl[4 * p + 3] = r_lb;
  for (int d = 0; d < 3; ++d) {
    a.kin[6 * p + d] = a.v[3 * p + d] / a.Uc * a.u_lb;      // get_vel_lb
    a.kin[6 * p + 3 + d] = a.w[3 * p + d] * a.tc * a.dt_d;  // get_freq_lb
  }
  int* z = a.zone + 6 * p;
  int* rz = a.rzone + 6 * p;
  for (int d = 0; d < 3; ++d) {
    int lo = (int)(loc[d] - r_lb) - 2;
    int hi = (int)(loc[d] + r_lb) + 2;
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
  if (zone_empty(z)) return;
  int b[6];
  brick_range(z, a.bi, a.bj, a.bk, b);
  for (int x = b[0]; x <= b[1]; ++x)
    for (int y = b[2]; y <= b[3]; ++y)
      for (int w = b[4]; w <= b[5]; ++w)
        atomicAdd(&a.count[((int64_t)x * a.nbj + y) * a.nbk + w], 1);
}

// Exclusive scan of brick counts: 1024 per block, then block sums.
__global__ __launch_bounds__(1024) void k_scan_block(const int* count, int* offset, int* sums,
                                                     int64_t n) {
  __shared__ int s[1024];
  const int64_t t = (int64_t)blockIdx.x * 1024 + threadIdx.x;
  const int v = t < n ? count[t] : 0;
  s[threadIdx.x] = v;
  __syncthreads();
  for (int d = 1; d < 1024; d <<= 1) {
    int add = threadIdx.x >= d ? s[threadIdx.x - d] : 0;
    __syncthreads();
    s[threadIdx.x] += add;
    __syncthreads();
  }
  if (t < n) offset[t] = s[threadIdx.x] - v;
  if (threadIdx.x == 1023) sums[blockIdx.x] = s[1023];
}

__global__ __launch_bounds__(1024) void k_scan_sums(int* sums, int nb) {
  __shared__ int s[1024];
  // sequential carry over chunks of 1024 block sums
  int carry = 0;
  for (int base = 0; base < nb; base += 1024) {
    const int t = base + threadIdx.x;
    const int v = t < nb ? sums[t] : 0;
    s[threadIdx.x] = v;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {
      int add = threadIdx.x >= d ? s[threadIdx.x - d] : 0;
      __syncthreads();
      s[threadIdx.x] += add;
      __syncthreads();
    }
    if (t < nb) sums[t] = carry + s[threadIdx.x] - v;
    const int tot = s[1023];
    __syncthreads();
    carry += tot;
  }
}

__global__ __launch_bounds__(1024) void k_scan_add(const int* count, int* offset,
                                                   const int* sums, int* cursor,
                                                   int* nonempty, int* nonempty_n, int64_t n) {
  const int64_t t = (int64_t)blockIdx.x * 1024 + threadIdx.x;
  if (t >= n) return;
  offset[t] += sums[blockIdx.x];
  cursor[t] = 0;
  if (count[t] > 0) {
    const int slot = atomicAdd(nonempty_n, 1);
    nonempty[slot] = (int)t;
  }
}

__global__ __launch_bounds__(256) void k_fill_lists(const PrepArgs a, const int* offset,
                                                    int* cursor, int* list) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= a.n) return;
  const int* z = a.zone + 6 * p;
  if (zone_empty(z)) return;
  int b[6];
  brick_range(z, a.bi, a.bj, a.bk, b);
  for (int x = b[0]; x <= b[1]; ++x)
    for (int y = b[2]; y <= b[3]; ++y)
      for (int w = b[4]; w <= b[5]; ++w) {
        const int64_t id = ((int64_t)x * a.nbj + y) * a.nbk + w;
        const int pos = atomicAdd(&cursor[id], 1);
        list[offset[id] + pos] = p;
      }
}

// One workgroup (256 threads = one brick of 256 nodes) per touched brick.
__global__ __launch_bounds__(256) void k_map(const MapArgs a, int* overflow) {
  __shared__ int raw[kMapMaxList];
  __shared__ int sorted[kMapMaxList];
  const Geom& g = a.g;
  const int64_t plane = g.plane;
  const int nne = *a.nonempty_n;
  for (int e = blockIdx.x; e < nne; e += gridDim.x) {
    const int b = a.nonempty[e];
    const int m = a.count[b];
    const int base = a.offset[b];
    if (m > kMapMaxList) {
      if (threadIdx.x == 0) atomicOr(overflow, 1);
      continue;  // uniform across the workgroup
    }
    for (int t = threadIdx.x; t < m; t += 256) raw[t] = a.list[base + t];
    __syncthreads();
    // entries are distinct, so rank = number of smaller entries
    for (int t = threadIdx.x; t < m; t += 256) {
      const int v = raw[t];
      int r = 0;
      for (int u = 0; u < m; ++u) r += raw[u] < v;
      sorted[r] = v;
    }
    __syncthreads();

    const int bx = b / (a.nbj * a.nbk);
    const int by = (b / a.nbk) % a.nbj;
    const int bz = b % a.nbk;
    const int lk = threadIdx.x % a.bk;
    const int lj = (threadIdx.x / a.bk) % a.bj;
    const int li = threadIdx.x / (a.bk * a.bj);
    const int i = bx * a.bi + li, j = by * a.bj + lj, k = bz * a.bk + lk;
    const bool valid = i < g.sx && j < g.sy && k < g.sz;
    if (valid) {
      const int n = (i * g.sy + j) * g.sz + k;
      int t0 = a.s.tag[n], t1 = a.s.tag[plane + n];
      double sf0 = a.s.sf[n], sf1 = a.s.sf[plane + n];
      double u0[3] = {a.s.us[n], a.s.us[plane + n], a.s.us[2 * plane + n]};
      double u1[3] = {a.s.us[3 * plane + n], a.s.us[4 * plane + n], a.s.us[5 * plane + n]};
      bool d0 = false, d1 = false;
      // reference (halo'd) node indices, as used in the geometry
      const int ri = i + g.off[0], rj = j + g.off[1], rk = k + g.off[2];
      for (int t = 0; t < m; ++t) {
        const int p = sorted[t];
        const int* z = a.zone + 6 * p;
        if (i < z[0] || i >= z[1] || j < z[2] || j >= z[3] || k < z[4] || k >= z[5]) continue;
        const double xL = a.xl[4 * p + 0], yL = a.xl[4 * p + 1], zL = a.xl[4 * p + 2];
        const double rL = a.xl[4 * p + 3];
        const int tag = a.tag[p];
        const double sf = solid_fraction(g.dim, a.hsq_half, ri, rj, rk, xL, yL, zL, rL);
        int id_old = 0;
        double sf_old = 0.0;
        if (tag == t0) {
          id_old = t0;
          sf_old = sf0;
        } else if (tag == t1) {
          id_old = t1;
          sf_old = sf1;
        }
        const int dec = (sf > 0.00001) + 2 * (sf_old > 0.00001);
        const double dx = (double)ri - xL;
        const double dy = (double)rj - yL;
        const double dz = (double)rk - zL;
        const double* kn = a.kin + 6 * p;
        double un[3];
        un[0] = kn[0] + kn[4] * dz - kn[5] * dy;
        un[1] = kn[1] + kn[5] * dx - kn[3] * dz;
        un[2] = kn[2] + kn[3] * dy - kn[4] * dx;
        bool do_set = false, do_clear = false;
        switch (dec) {
          case 0: do_clear = true; break;
          case 1: do_set = true; break;
          case 2: do_clear = (id_old == tag); break;
          case 3: do_set = (sf > sf_old || id_old == tag); break;
        }
        if (do_set) {  // setParticleOnLattice, src/LBM_PSM_lattice.cpp:257-299
          int s = -1;
          if (t0 == tag) s = 0;
          else if (t1 == tag) s = 1;
          else if (t0 == 0) s = 0;
          else if (t1 == 0) s = 1;
          if (s == 0) {
            t0 = tag; sf0 = sf; u0[0] = un[0]; u0[1] = un[1]; u0[2] = un[2]; d0 = true;
          } else if (s == 1) {
            t1 = tag; sf1 = sf; u1[0] = un[0]; u1[1] = un[1]; u1[2] = un[2]; d1 = true;
          }
        } else if (do_clear) {  // setToZero, src/LBM_PSM_lattice.cpp:302-324
          if (t0 == tag) {
            t0 = 0; sf0 = 0.0; u0[0] = u0[1] = u0[2] = 0.0; d0 = true;
          } else if (t1 == tag) {
            t1 = 0; sf1 = 0.0; u1[0] = u1[1] = u1[2] = 0.0; d1 = true;
          }
        }
      }
      if (d0) {
        a.s.tag[n] = t0;
        a.s.sf[n] = sf0;
        a.s.us[n] = u0[0];
        a.s.us[plane + n] = u0[1];
        a.s.us[2 * plane + n] = u0[2];
        a.s.Fh[n] = 0.0;
        a.s.Fh[plane + n] = 0.0;
        a.s.Fh[2 * plane + n] = 0.0;
      }
      if (d1) {
        a.s.tag[plane + n] = t1;
        a.s.sf[plane + n] = sf1;
        a.s.us[3 * plane + n] = u1[0];
        a.s.us[4 * plane + n] = u1[1];
        a.s.us[5 * plane + n] = u1[2];
        a.s.Fh[3 * plane + n] = 0.0;
        a.s.Fh[4 * plane + n] = 0.0;
        a.s.Fh[5 * plane + n] = 0.0;
      }
      if (d0 || d1) a.s.mask[n] = (uint8_t)((t0 > 0 ? 1 : 0) | (t1 > 0 ? 2 : 0));
    }
    __syncthreads();  // LDS lists are rewritten by the next brick
  }
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
  const int64_t tot = (ni > 0 && nj > 0 && nk > 0) ? (int64_t)ni * nj * nk : 0;
  for (int64_t t = threadIdx.x; t < tot; t += 256) {
    const int lk = (int)(t % nk);
    const int lj = (int)((t / nk) % nj);
    const int li = (int)(t / ((int64_t)nk * nj));
    const int i = z[0] + li, j = z[2] + lj, k = z[4] + lk;
    const int n = (i * g.sy + j) * g.sz + k;
    int s;
    if (tag == a.slot_tag[n]) s = 0;
    else if (tag == a.slot_tag[plane + n]) s = 1;
    else continue;
    const double F0 = a.Fh[(3 * s) * plane + n];
    const double F1 = a.Fh[(3 * s + 1) * plane + n];
    const double F2 = a.Fh[(3 * s + 2) * plane + n];
    acc[0] += F0 * ff;
    acc[1] += F1 * ff;
    acc[2] += F2 * ff;
    const double dx = (double)(i + g.off[0]) - xL;
    const double dy = (double)(j + g.off[1]) - yL;
    const double dz = (double)(k + g.off[2]) - zL;
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

hipError_t launch_prep_count(const PrepArgs& a, hipStream_t s) {
  if (a.n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_prep_count, grid_for(a.n, 256), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_scan_bricks(BrickDev& b, hipStream_t s) {
  const int64_t n = b.nbricks;
  const int nb = (int)((n + 1023) / 1024);
  hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(256), 0, s, b.nonempty_n, (int64_t)1, 0);
  hipLaunchKernelGGL(k_scan_block, dim3(nb), dim3(1024), 0, s, b.count, b.offset, b.block_sums, n);
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, s, b.block_sums, nb);
  hipLaunchKernelGGL(k_scan_add, dim3(nb), dim3(1024), 0, s, b.count, b.offset, b.block_sums,
                     b.cursor, b.nonempty, b.nonempty_n, n);
  return hipGetLastError();
}

hipError_t launch_fill_lists(const PrepArgs& a, const BrickDev& b, hipStream_t s) {
  if (a.n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_lists, grid_for(a.n, 256), dim3(256), 0, s, a, b.offset, b.cursor,
                     b.list);
  return hipGetLastError();
}

hipError_t launch_map(const MapArgs& a, int* overflow, hipStream_t s, int grid) {
  hipLaunchKernelGGL(k_map, dim3(grid), dim3(256), 0, s, a, overflow);
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
