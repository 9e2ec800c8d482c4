// Discrete adjoint of the FP64 LBM-PSM step on a fully periodic lattice
// without Zou-He faces (no counterpart in the reference).
//
// Forward model of one step t, per node x (the FP64 collide kernel's):
//   g_q      = h^{t-1}_q(x - e_q)          (pull; in place for the initial
//                                           post-stream state)
//   rho      = sum g,  u = (sum g e + F/2) / rho
//   feq_q    = rho w_q P_q(u),  fs_{s,q} = rho w_q P_q(u_s),
//   P_q(v)   = 1 + 3 e.v + 4.5 (e.v)^2 - 1.5 v.v
//   sc_{s,q} = fs_{s,q} - g_q + (1 - om)(g_q - feq_q)
//   h^t_q(x) = g_q + (1 - Btot) om (feq_q - g_q) + sum_s B_s sc_{s,q}
//              + (1 - om/2) Fi_q(u, F)       (Guo, when |F| > 0)
//   Fh_s    += -B_s sc_{s,q} e_q            (node force on slot s)
// with om = 1/tau and B_s the (renormalised) solid fractions.  The
// particles are frozen: slots, B_s and the lever arms are the same every
// step, u_s = v_lb + w_lb x r depends on the particle velocities only.
//
// The loss depends on the per-entry force, torque and stresslet of the last
// step, which are linear in that step's node forces.  Its gradient is seeded
// per node and slot (k_adj_seed) and carried backwards:
//   mu_q(x)      = lambda^{t+1}_q(x + e_q)  (adjoint of the pull)
//   lambda^t(x)  = J_x^T mu(x)              (adjoint of the node update)
// while the parameter gradients are accumulated per node:
//   d/d om, d/dF (all nodes), d/du_s (occupied slots).
// Entries then collect d/dv and d/dw over the nodes whose slot they wrote
// (k_adj_writer, k_adj_entry), and d/d om and d/dF are summed in a fixed
// order (k_adj_param).
#include "lbmpsm_internal.h"

namespace lbmpsm {

__device__ __forceinline__ int wrapc(int c, int n) {
  return c < 0 ? c + n : (c >= n ? c - n : c);
}

template <int Q>
__global__ __launch_bounds__(256) void k_adj_step(const AdjArgs a) {
  const Geom& g = a.g;
  const int n = (int)blockIdx.x * 256 + (int)threadIdx.x;
  if (n >= (int)g.nodes) return;
  const int k = n % g.sz;
  const int rj = n / g.sz;
  const int j = rj % g.sy;
  const int i = rj / g.sy;
  const int64_t P = g.plane;
  const double omega = a.omega, inv2 = a.invCs2, inv4 = a.invCs4;
  const double F0 = a.F[0], F1 = a.F[1], F2 = a.F[2];

  double f[Q], mu[Q];
  Unroll<0, Q>::run([&]<int q>() {
    constexpr int ex = VelSet<Q>::ex[q], ey = VelSet<Q>::ey[q], ez = VelSet<Q>::ez[q];
    const int src = (wrapc(i - ex, g.sx) * g.sy + wrapc(j - ey, g.sy)) * g.sz + wrapc(k - ez, g.sz);
    const int dst = (wrapc(i + ex, g.sx) * g.sy + wrapc(j + ey, g.sy)) * g.sz + wrapc(k + ez, g.sz);
    f[q] = a.hprev[pop_index_f64<Q>(q, P, (unsigned)(a.inplace ? n : src))];
    mu[q] = a.lam_in ? a.lam_in[q * P + dst] : 0.0;
  });

  double rho, jx, jy, jz;
  moments<Q, double>(f, rho, jx, jy, jz);
  jx = jx + F0 * 0.5;
  jy = jy + F1 * 0.5;
  jz = jz + F2 * 0.5;
  const double ux = jx / rho, uy = jy / rho, uz = jz / rho;
  const double usqt = 0.5 * (ux * ux + uy * uy + uz * uz) * inv2;

  const uint8_t m = a.mask[n];
  const bool o1 = m & 1, o2 = m & 2;
  double B1 = o1 ? a.sf[n] : 0.0;
  double B2 = o2 ? a.sf[P + n] : 0.0;
  double us1[3] = {0.0, 0.0, 0.0}, us2[3] = {0.0, 0.0, 0.0};
  double sd1[3] = {0.0, 0.0, 0.0}, sd2[3] = {0.0, 0.0, 0.0};  // node-force seeds
  if (o1)
    for (int d = 0; d < 3; ++d) {
      us1[d] = a.us[d * P + n];
      if (a.seed) sd1[d] = a.seed[d * P + n];
    }
  if (o2)
    for (int d = 0; d < 3; ++d) {
      us2[d] = a.us[(3 + d) * P + n];
      if (a.seed) sd2[d] = a.seed[(3 + d) * P + n];
    }
  const double Btot0 = B1 + B2;
  double Btot = Btot0;
  if (Btot0 > 1.0) {
    B1 = B1 / Btot0;
    B2 = B2 / Btot0;
    Btot = 1.0;
  }
  const double u1sq = 0.5 * (us1[0] * us1[0] + us1[1] * us1[1] + us1[2] * us1[2]) * inv2;
  const double u2sq = 0.5 * (us2[0] * us2[0] + us2[1] * us2[1] + us2[2] * us2[2]) * inv2;

  double fdir[Q];
  double rhobar = 0.0, ombar = 0.0;
  double ub[3] = {0.0, 0.0, 0.0}, Fb[3] = {0.0, 0.0, 0.0};
  double u1b[3] = {0.0, 0.0, 0.0}, u2b[3] = {0.0, 0.0, 0.0};
  const double cf = (1.0 - Btot) * omega;
  Unroll<0, Q>::run([&]<int q>() {
    constexpr double w = VelSet<Q>::w[q];
    constexpr double ex = VelSet<Q>::ex[q], ey = VelSet<Q>::ey[q], ez = VelSet<Q>::ez[q];
    const double eu = ex * ux + ey * uy + ez * uz;
    const double Pq = 1.0 + eu * inv2 + eu * eu * 0.5 * inv4 - usqt;
    const double feq = rho * w * Pq;
    const double e1 = ex * us1[0] + ey * us1[1] + ez * us1[2];
    const double e2 = ex * us2[0] + ey * us2[1] + ez * us2[2];
    const double P1 = 1.0 + e1 * inv2 + 0.5 * e1 * e1 * inv4 - u1sq;
    const double P2 = 1.0 + e2 * inv2 + 0.5 * e2 * e2 * inv4 - u2sq;
    const double eF = ex * F0 + ey * F1 + ez * F2;
    const double Fi = w * (((ex - ux) * F0 + (ey - uy) * F1 + (ez - uz) * F2) * inv2 + eu * inv4 * eF);
    const double mq = mu[q];
    // adjoints of sc_1, sc_2 (update term and, on the last step, node force)
    const double s1b = o1 ? B1 * mq - B1 * (ex * sd1[0] + ey * sd1[1] + ez * sd1[2]) : 0.0;
    const double s2b = o2 ? B2 * mq - B2 * (ex * sd2[0] + ey * sd2[1] + ez * sd2[2]) : 0.0;
    const double sb = s1b + s2b;
    fdir[q] = mq * (1.0 - cf) - omega * sb;
    const double feqb = mq * cf - (1.0 - omega) * sb;
    ombar += mq * (1.0 - Btot) * (feq - f[q]) - sb * (f[q] - feq) - 0.5 * mq * Fi;
    const double Fib = mq * (1.0 - 0.5 * omega);
    rhobar += feqb * w * Pq + s1b * w * P1 + s2b * w * P2;
    const double cu = feqb * rho * w;
    const double e3[3] = {ex, ey, ez};
    const double uu[3] = {ux, uy, uz};
    const double FF[3] = {F0, F1, F2};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      ub[d] += cu * (e3[d] * inv2 + eu * e3[d] * inv4 - uu[d] * inv2) +
               Fib * w * (-FF[d] * inv2 + e3[d] * inv4 * eF);
      Fb[d] += Fib * w * ((e3[d] - uu[d]) * inv2 + eu * inv4 * e3[d]);
      u1b[d] += s1b * rho * w * (e3[d] * inv2 + e1 * e3[d] * inv4 - us1[d] * inv2);
      u2b[d] += s2b * rho * w * (e3[d] * inv2 + e2 * e3[d] * inv4 - us2[d] * inv2);
    }
  });
  // u = (j + F/2) / rho
  const double irho = 1.0 / rho;
  Fb[0] += 0.5 * ub[0] * irho;
  Fb[1] += 0.5 * ub[1] * irho;
  Fb[2] += 0.5 * ub[2] * irho;
  const double base = rhobar - (ub[0] * ux + ub[1] * uy + ub[2] * uz) * irho;
  Unroll<0, Q>::run([&]<int q>() {
    constexpr double ex = VelSet<Q>::ex[q], ey = VelSet<Q>::ey[q], ez = VelSet<Q>::ez[q];
    a.lam_out[q * P + n] = fdir[q] + base + (ub[0] * ex + ub[1] * ey + ub[2] * ez) * irho;
  });

  // accumulators (zeroed before the reverse sweep): om, F, u_s slot 0, u_s slot 1
  double* acc = a.acc;
  acc[n] += ombar;
  for (int d = 0; d < 3; ++d) acc[(1 + d) * P + n] += Fb[d];
  if (o1)
    for (int d = 0; d < 3; ++d) acc[(4 + d) * P + n] += u1b[d];
  if (o2)
    for (int d = 0; d < 3; ++d) acc[(7 + d) * P + n] += u2b[d];
}

// Gradient of the loss with respect to each node force of the last step:
// every entry whose reduction zone holds the node and whose tag is in one
// of its slots contributes, with the factors and lever arm of the hydro
// reduction (force ff F, torque tf r x F, stresslet -tf sym(r F)).
__global__ __launch_bounds__(256) void k_adj_seed(const AdjSeedArgs a) {
  const Geom& g = a.g;
  const int n = (int)blockIdx.x * 256 + (int)threadIdx.x;
  if (n >= (int)g.nodes) return;
  const int64_t P = g.plane;
  double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  const int t0 = a.slot_tag[n], t1 = a.slot_tag[P + n];
  if (a.mask[n]) {
    const int k = n % g.sz;
    const int rj = n / g.sz;
    const int j = rj % g.sy;
    const int i = rj / g.sy;
    const double ff = a.forceFactor, tf = a.torqueFactor;
    for (int e = 0; e < a.n; ++e) {
      const int te = a.tag[e];
      const int s = te == t0 ? 0 : (te == t1 ? 1 : -1);
      if (s < 0) continue;
      const int* z = a.rzone + 6 * e;
      if (i < z[0] || i >= z[1] || j < z[2] || j >= z[3] || k < z[4] || k >= z[5]) continue;
      const double dx = (double)(i + g.off[0]) - a.xl[4 * e + 0];
      const double dy = (double)(j + g.off[1]) - a.xl[4 * e + 1];
      const double dz = (double)(k + g.off[2]) - a.xl[4 * e + 2];
      const double* B = a.bar + 12 * e;  // F(3) T(3) S(6)
      acc[3 * s + 0] += ff * B[0] + tf * (B[4] * dz - B[5] * dy) -
                        tf * (B[6] * dx + 0.5 * B[9] * dy + 0.5 * B[10] * dz);
      acc[3 * s + 1] += ff * B[1] + tf * (B[5] * dx - B[3] * dz) -
                        tf * (B[7] * dy + 0.5 * B[9] * dx + 0.5 * B[11] * dz);
      acc[3 * s + 2] += ff * B[2] + tf * (B[3] * dy - B[4] * dx) -
                        tf * (B[8] * dz + 0.5 * B[10] * dx + 0.5 * B[11] * dy);
    }
  }
  for (int c = 0; c < 6; ++c) a.seed[c * P + n] = acc[c];
}

// Entry that wrote each occupied slot: with frozen particles the last entry
// in list order whose map zone holds the node and whose tag the slot holds
// (an entry of that tag with zero solid fraction would have cleared it).
__global__ __launch_bounds__(256) void k_adj_writer(const AdjSeedArgs a) {
  const Geom& g = a.g;
  const int n = (int)blockIdx.x * 256 + (int)threadIdx.x;
  if (n >= (int)g.nodes) return;
  const int64_t P = g.plane;
  int w0 = -1, w1 = -1;
  const uint8_t m = a.mask[n];
  if (m) {
    const int t0 = a.slot_tag[n], t1 = a.slot_tag[P + n];
    const int k = n % g.sz;
    const int rj = n / g.sz;
    const int j = rj % g.sy;
    const int i = rj / g.sy;
    for (int e = 0; e < a.n; ++e) {
      const int te = a.tag[e];
      if (te != t0 && te != t1) continue;
      const int* z = a.zone + 6 * e;
      if (i < z[0] || i >= z[1] || j < z[2] || j >= z[3] || k < z[4] || k >= z[5]) continue;
      if ((m & 1) && te == t0) w0 = e;
      if ((m & 2) && te == t1) w1 = e;
    }
  }
  a.writer[n] = w0;
  a.writer[P + n] = w1;
}

// d/dv_lb and d/dw_lb of one entry: sums over the nodes of its map zone
// whose slot it wrote, u_s = v + w x r  =>  dw = r x du_s.
__global__ __launch_bounds__(256) void k_adj_entry(const AdjSeedArgs a, const double* acc,
                                                   double* out) {
  const int e = blockIdx.x;
  if (e >= a.n) return;
  const Geom& g = a.g;
  const int64_t P = g.plane;
  const int* z = a.zone + 6 * e;
  const int ni = z[1] - z[0], nj = z[3] - z[2], nk = z[5] - z[4];
  const int tot = (ni > 0 && nj > 0 && nk > 0) ? ni * nj * nk : 0;
  const int njk = nj * nk;
  double s[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int t = threadIdx.x; t < tot; t += 256) {
    const int li = t / njk;
    const int rem = t - li * njk;
    const int lj = rem / nk;
    const int lk = rem - lj * nk;
    const int i = z[0] + li, j = z[2] + lj, k = z[4] + lk;
    const int n = (i * g.sy + j) * g.sz + k;
    for (int sl = 0; sl < 2; ++sl) {
      if (a.writer[sl * P + n] != e) continue;
      const double b0 = acc[(4 + 3 * sl) * P + n];
      const double b1 = acc[(5 + 3 * sl) * P + n];
      const double b2 = acc[(6 + 3 * sl) * P + n];
      const double dx = (double)(i + g.off[0]) - a.xl[4 * e + 0];
      const double dy = (double)(j + g.off[1]) - a.xl[4 * e + 1];
      const double dz = (double)(k + g.off[2]) - a.xl[4 * e + 2];
      s[0] += b0;
      s[1] += b1;
      s[2] += b2;
      s[3] += dy * b2 - dz * b1;
      s[4] += dz * b0 - dx * b2;
      s[5] += dx * b1 - dy * b0;
    }
  }
  __shared__ double red[6][256];
  for (int c = 0; c < 6; ++c) red[c][threadIdx.x] = s[c];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int c = 0; c < 6; ++c) red[c][threadIdx.x] += red[c][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < 6) out[6 * e + threadIdx.x] = red[threadIdx.x][0];
}

// Per-block partial sums of d/d om and d/dF over all nodes.
__global__ __launch_bounds__(256) void k_adj_param(const Geom g, const double* acc,
                                                   double* partials) {
  double s[4] = {0.0, 0.0, 0.0, 0.0};
  for (int64_t n = (int64_t)blockIdx.x * 256 + threadIdx.x; n < g.nodes;
       n += (int64_t)gridDim.x * 256)
    for (int c = 0; c < 4; ++c) s[c] += acc[c * g.plane + n];
  __shared__ double red[4][256];
  for (int c = 0; c < 4; ++c) red[c][threadIdx.x] = s[c];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int c = 0; c < 4; ++c) red[c][threadIdx.x] += red[c][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < 4) partials[4 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

static inline dim3 node_grid(const Geom& g) { return dim3((unsigned)((g.nodes + 255) / 256)); }

hipError_t launch_adj_step(const AdjArgs& a, hipStream_t s) {
  if (a.g.q == 19) hipLaunchKernelGGL(k_adj_step<19>, node_grid(a.g), dim3(256), 0, s, a);
  else hipLaunchKernelGGL(k_adj_step<9>, node_grid(a.g), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_adj_seed(const AdjSeedArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_adj_seed, node_grid(a.g), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_adj_writer(const AdjSeedArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_adj_writer, node_grid(a.g), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_adj_entry(const AdjSeedArgs& a, const double* acc, double* out,
                            hipStream_t s) {
  if (a.n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_adj_entry, dim3((unsigned)a.n), dim3(256), 0, s, a, acc, out);
  return hipGetLastError();
}

hipError_t launch_adj_param(const Geom& g, const double* acc, double* partials, int nblocks,
                            hipStream_t s) {
  hipLaunchKernelGGL(k_adj_param, dim3((unsigned)nblocks), dim3(256), 0, s, g, acc, partials);
  return hipGetLastError();
}

}  // namespace lbmpsm
