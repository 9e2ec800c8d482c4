/*
 * CPU oracle for the LBM-PSM step — TEST INFRASTRUCTURE ONLY (see header).
 *
 * Every function names the reference lines it restates.  Floating-point
 * expressions keep the reference's operand order so that, compiled with
 * -ffp-contract=off, each result is the same double the reference computes.
 */
#include "lbmpsm_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Velocity sets, src/LBM_PSM_lattice.h:186-209. */
static const int D2Q9_ex[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
static const int D2Q9_ey[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
static const int D2Q9_ez[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
static const double D2Q9_w[9] = {4.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0,
                                 1.0 / 9.0,  1.0 / 9.0,  1.0 / 36.0,
                                 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};
static const int D3Q19_ex[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0};
static const int D3Q19_ey[19] = {0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1};
static const int D3Q19_ez[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, -1, 1, -1, 1};
static const double D3Q19_w[19] = {
    1.0 / 3.0,  1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0,
    1.0 / 18.0, 1.0 / 18.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
    1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
    1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};

struct lbo_lattice {
  int dim, q;
  int nx, ny, nz; /* halo'd local sizes */
  int periodic[3];
  double tau, omega;
  double F[3];
  double Fmag2;
  const int *ex, *ey, *ez;
  const double* w;
  double csPow2, csPow4, invCsPow2, invCsPow4;
  double *f_curr, *f_next;
  double *rho, *u;
  int* pid;       /* [node*2+s] particleID        */
  double* sf;     /* [node*2+s] solidFraction     */
  double* us;     /* [node*6+3s+d] particleVelocity */
  double* Fh;     /* [node*6+3s+d] hydrodynamicForce */
  /* per axis, the lattice offset the map and the reduction subtract from a
   * particle's global lattice position: (nTotal - p*2*envelopeWidth - 1),
   * src/LBM_PSM_exchangeParticleData.cpp:25-35, 176-183; -1 on one rank */
  double roff[3];
};

/* src/LBM_PSM_unitConversion.cpp:15-38 */
void lbo_units_init(lbo_units* U, double rhof_, double nu_, double lc_, double Re_,
                    int N_, double tau_, int dim_) {
  U->rhof = rhof_; U->nu = nu_; U->lc = lc_; U->Re = Re_; U->N = N_;
  U->tau = tau_; U->dim = dim_;
  U->Uc_d = 1.0;
  U->Lc_d = 1.0;
  U->cs = 1.0 / sqrt(3.0);
  U->csPow2 = 1.0 / 3.0;
  U->nu_lb = (tau_ - 0.5) * U->csPow2;
  U->dx = lc_ / (double)(N_ - 1);
  U->dx_d = U->Lc_d / (double)(N_ - 1);
  U->Uc = Re_ * nu_ / lc_;
  U->u_lb = Re_ * U->nu_lb / (double)(N_ - 1);
  U->dt_d = U->dx_d * U->u_lb / U->Uc_d;
  U->tc = lc_ / U->Uc;
  U->forceFactor = rhof_ * pow(lc_, 4) / pow(U->tc, 2) * pow(U->dx_d, 4) / pow(U->dt_d, 2);
  U->torqueFactor = rhof_ * pow(lc_, 5) / pow(U->tc, 2) * pow(U->dx_d, 5) / pow(U->dt_d, 2);
  U->volumeForceFactor = rhof_ * lc_ / pow(U->tc, 2) * U->dx_d / pow(U->dt_d, 2);
}

static inline int idx1(const lbo_lattice* L, int i, int j, int k) {
  return i * L->ny * L->nz + j * L->nz + k; /* src/LBM_PSM_lattice.h:88 */
}

/* src/LBM_PSM_lattice.cpp:16-118 (one rank) and 124-140 */
lbo_lattice* lbo_create(int dim, int nx, int ny, int nz, const int periodic[3],
                        double tau, const double F_lbm[3]) {
  if (dim != 2 && dim != 3) return NULL;
  if (nx < 2 || ny < 2 || (dim == 3 && nz < 2)) return NULL;
  lbo_lattice* L = (lbo_lattice*)calloc(1, sizeof(lbo_lattice));
  if (!L) return NULL;
  L->dim = dim;
  L->nx = nx + 2;
  L->ny = ny + 2;
  L->nz = (dim == 3) ? nz + 2 : 1;
  for (int d = 0; d < 3; ++d) L->periodic[d] = periodic[d] ? 1 : 0;
  if (dim == 2) {
    L->q = 9; L->ex = D2Q9_ex; L->ey = D2Q9_ey; L->ez = D2Q9_ez; L->w = D2Q9_w;
  } else {
    L->q = 19; L->ex = D3Q19_ex; L->ey = D3Q19_ey; L->ez = D3Q19_ez; L->w = D3Q19_w;
  }
  L->csPow2 = 1.0 / 3.0;
  L->csPow4 = L->csPow2 * L->csPow2;
  L->invCsPow2 = 1.0 / L->csPow2;
  L->invCsPow4 = 1.0 / L->csPow4;
  L->tau = tau;
  L->omega = 1.0 / tau; /* src/LBM_PSM_BGK_dynamics.cpp:16 */
  for (int d = 0; d < 3; ++d) L->F[d] = F_lbm[d];
  for (int d = 0; d < 3; ++d) L->roff[d] = (0 - ((double)0 * 2.0 * (double)1) - 1);
  L->Fmag2 = 0.0;
  size_t nodes = (size_t)L->nx * L->ny * L->nz;
  L->f_curr = (double*)calloc(nodes * L->q, sizeof(double));
  L->f_next = (double*)calloc(nodes * L->q, sizeof(double));
  L->rho = (double*)calloc(nodes, sizeof(double));
  L->u = (double*)calloc(nodes * 3, sizeof(double));
  L->pid = (int*)calloc(nodes * 2, sizeof(int));
  L->sf = (double*)calloc(nodes * 2, sizeof(double));
  L->us = (double*)calloc(nodes * 6, sizeof(double));
  L->Fh = (double*)calloc(nodes * 6, sizeof(double));
  if (!L->f_curr || !L->f_next || !L->rho || !L->u || !L->pid || !L->sf || !L->us || !L->Fh) {
    lbo_destroy(L);
    return NULL;
  }
  return L;
}

void lbo_destroy(lbo_lattice* L) {
  if (!L) return;
  free(L->f_curr); free(L->f_next); free(L->rho); free(L->u);
  free(L->pid); free(L->sf); free(L->us); free(L->Fh);
  free(L);
}

void lbo_dims(const lbo_lattice* L, int out[4]) {
  out[0] = L->nx; out[1] = L->ny; out[2] = L->nz; out[3] = L->q;
}

double* lbo_f(lbo_lattice* L) { return L->f_curr; }
double* lbo_rho(lbo_lattice* L) { return L->rho; }
double* lbo_u(lbo_lattice* L) { return L->u; }
int* lbo_slot_tag(lbo_lattice* L) { return L->pid; }
double* lbo_slot_sf(lbo_lattice* L) { return L->sf; }
double* lbo_slot_u(lbo_lattice* L) { return L->us; }
double* lbo_slot_F(lbo_lattice* L) { return L->Fh; }

int lbo_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void lbo_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

/* feq from the node's stored rho and u, src/LBM_PSM_dynamics.h:43-59 */
static inline double feq_node(const lbo_lattice* L, int iq, int n1) {
  const double ex_i = L->ex[iq];
  const double ey_i = L->ey[iq];
  const double ez_i = L->ez[iq];
  const double w_i = L->w[iq];
  const double ux = L->u[n1 * 3];
  const double uy = L->u[n1 * 3 + 1];
  const double uz = L->u[n1 * 3 + 2];
  double eiq_dot_u = ex_i * ux + ey_i * uy + ez_i * uz;
  return L->rho[n1] * w_i *
         (1.0 + eiq_dot_u * L->invCsPow2 + eiq_dot_u * eiq_dot_u * 0.5 * L->invCsPow4 -
          0.5 * (ux * ux + uy * uy + uz * uz) * L->invCsPow2);
}

/* feq for an explicit rho and velocity, src/LBM_PSM_dynamics.h:72-85 */
static inline double feq_vel(const lbo_lattice* L, int iq, double rho, const double* u) {
  const double ex_i = L->ex[iq];
  const double ey_i = L->ey[iq];
  const double ez_i = L->ez[iq];
  const double eiq_dot_u = ex_i * u[0] + ey_i * u[1] + ez_i * u[2];
  const double u_dot_u = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
  return rho * L->w[iq] *
         (1.0 + eiq_dot_u * L->invCsPow2 + 0.5 * eiq_dot_u * eiq_dot_u * L->invCsPow4 -
          0.5 * u_dot_u * L->invCsPow2);
}

/* Guo forcing term, src/LBM_PSM_dynamics.h:102-116 */
static inline double F_iq(const lbo_lattice* L, int iq, int n1) {
  double ex_i = L->ex[iq];
  double ey_i = L->ey[iq];
  double ez_i = L->ez[iq];
  double w_i = L->w[iq];
  double ux = L->u[n1 * 3];
  double uy = L->u[n1 * 3 + 1];
  double uz = L->u[n1 * 3 + 2];
  const double* F_ = L->F;
  return w_i * (((ex_i - ux) * F_[0] + (ey_i - uy) * F_[1] + (ez_i - uz) * F_[2]) * L->invCsPow2 +
                ((ex_i * ux + ey_i * uy + ez_i * uz) * L->invCsPow4 *
                 (ex_i * F_[0] + ey_i * F_[1] + ez_i * F_[2])));
}

/* src/LBM_PSM_BGK_dynamics.cpp:22-44 plus the slot reset of
 * src/LBM_PSM_lattice.cpp:142-186 */
void lbo_initialise(lbo_lattice* L, double rho_, double ux_, double uy_, double uz_) {
  int q = L->q;
  for (int i = 0; i < L->nx; ++i)
    for (int j = 0; j < L->ny; ++j)
      for (int k = 0; k < L->nz; ++k) {
        int n1 = idx1(L, i, j, k);
        L->rho[n1] = rho_;
        L->u[n1 * 3] = ux_;
        L->u[n1 * 3 + 1] = uy_;
        L->u[n1 * 3 + 2] = uz_;
        for (int iq = 0; iq < q; ++iq) {
          double feq_val = feq_node(L, iq, n1);
          L->f_curr[n1 * q + iq] = feq_val;
          L->f_next[n1 * q + iq] = feq_val;
        }
        for (int s = 0; s < 2; ++s) {
          L->pid[n1 * 2 + s] = 0;
          L->sf[n1 * 2 + s] = 0.0;
        }
        for (int c = 0; c < 6; ++c) {
          L->us[n1 * 6 + c] = 0.0;
          L->Fh[n1 * 6 + c] = 0.0;
        }
      }
  L->Fmag2 = L->F[0] * L->F[0] + L->F[1] * L->F[1] + L->F[2] * L->F[2];
}

/* setParticleOnLattice, src/LBM_PSM_lattice.cpp:257-299 */
static void slot_set(lbo_lattice* L, int n1, int pID, const double uP[3], double eps) {
  int* id = L->pid + n1 * 2;
  int s;
  if (id[0] == pID) s = 0;
  else if (id[1] == pID) s = 1;
  else if (id[0] == 0) s = 0;
  else if (id[1] == 0) s = 1;
  else return; /* both slots held by other particles: dropped */
  id[s] = pID;
  L->sf[n1 * 2 + s] = eps;
  for (int d = 0; d < 3; ++d) {
    L->us[n1 * 6 + 3 * s + d] = uP[d];
    L->Fh[n1 * 6 + 3 * s + d] = 0.0;
  }
}

/* setToZero, src/LBM_PSM_lattice.cpp:302-324 */
static void slot_clear(lbo_lattice* L, int n1, int pID) {
  int* id = L->pid + n1 * 2;
  int s;
  if (id[0] == pID) s = 0;
  else if (id[1] == pID) s = 1;
  else return;
  id[s] = 0;
  L->sf[n1 * 2 + s] = 0.0;
  for (int d = 0; d < 3; ++d) {
    L->us[n1 * 6 + 3 * s + d] = 0.0;
    L->Fh[n1 * 6 + 3 * s + d] = 0.0;
  }
}

/* calcSolidFraction, src/LBM_PSM_exchangeParticleData.cpp:116-169 */
double lbo_solid_fraction(const lbo_lattice* L, int i, int j, int k, double xP_LB,
                          double yP_LB, double zP_LB, double rP_LB) {
  const int slicesPerDim = 10;
  double sliceWidth = 1.0 / ((double)slicesPerDim);
  double fraction = 1.0 / ((double)(slicesPerDim * slicesPerDim));
  double dx_sq[10], dy_sq[10], dz_sq[10];
  if (L->dim == 3) fraction = 1.0 / ((double)(slicesPerDim * slicesPerDim * slicesPerDim));
  double sqrt2half = sqrt(2.1) / 2.0;
  if (L->dim == 3) sqrt2half = sqrt(3.1) / 2.0;
  double dx = (double)i - xP_LB;
  double dy = (double)j - yP_LB;
  double dz = (double)k - zP_LB;
  double dist = dx * dx + dy * dy + dz * dz;
  double rP = rP_LB + sqrt2half;
  if (dist > rP * rP) return 0.0;
  double rM = rP_LB - sqrt2half;
  if (dist < rM * rM) return 1.0;
  double rSq = rP_LB * rP_LB;
  for (int a = 0; a < slicesPerDim; ++a) {
    double delta = -0.5 + ((double)a + 0.5) * sliceWidth;
    double dx_i = dx + delta; dx_sq[a] = dx_i * dx_i;
    double dy_i = dy + delta; dy_sq[a] = dy_i * dy_i;
    double dz_i = dz + delta; dz_sq[a] = dz_i * dz_i;
  }
  int n = 0;
  if (L->dim == 2) {
    for (int a = 0; a < slicesPerDim; ++a)
      for (int b = 0; b < slicesPerDim; ++b)
        if (dx_sq[a] + dy_sq[b] < rSq) ++n;
  } else {
    for (int a = 0; a < slicesPerDim; ++a)
      for (int b = 0; b < slicesPerDim; ++b)
        for (int c = 0; c < slicesPerDim; ++c)
          if (dx_sq[a] + dy_sq[b] + dz_sq[c] < rSq) ++n;
  }
  return fraction * ((double)n);
}

static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
static inline int maxi(int a, int b) { return a > b ? a : b; }

/* setParticlesOnLattice, src/LBM_PSM_exchangeParticleData.cpp:24-112, one rank:
 * the local offset (nTotal - coord*2*envelope - 1) is -1.0 on every axis. */
void lbo_set_particles(lbo_lattice* L, const lbo_units* U, const double origin[3], int n,
                       const int* tag, const double* x, const double* v,
                       const double* omega, const double* radius) {
  for (int p = 0; p < n; ++p) {
    double x_lb_global = (x[3 * p + 0] - origin[0]) / U->dx;
    double y_lb_global = (x[3 * p + 1] - origin[1]) / U->dx;
    double z_lb_global = (x[3 * p + 2] - origin[2]) / U->dx;
    double x_loc = x_lb_global - L->roff[0];
    double y_loc = y_lb_global - L->roff[1];
    double z_loc = z_lb_global - L->roff[2];
    double r_lb = radius[p] / U->dx;
    int zone[3][2] = {{(int)(x_loc - r_lb) - 2, (int)(x_loc + r_lb) + 2},
                      {(int)(y_loc - r_lb) - 2, (int)(y_loc + r_lb) + 2},
                      {(int)(z_loc - r_lb) - 2, (int)(z_loc + r_lb) + 2}};
    for (int e = 0; e < 2; ++e) {
      zone[0][e] = clampi(zone[0][e], 0, L->nx);
      zone[1][e] = clampi(zone[1][e], 0, L->ny);
      zone[2][e] = maxi(clampi(zone[2][e], 0, L->nz), e);
    }
    const double uTrans[3] = {v[3 * p + 0] / U->Uc * U->u_lb, v[3 * p + 1] / U->Uc * U->u_lb,
                              v[3 * p + 2] / U->Uc * U->u_lb};
    const double omegaLB[3] = {omega[3 * p + 0] * U->tc * U->dt_d,
                               omega[3 * p + 1] * U->tc * U->dt_d,
                               omega[3 * p + 2] * U->tc * U->dt_d};
    /* The nodes of one particle's zone are independent of each other (each
     * reads and writes only its own record), so they are split over threads;
     * particles stay in list order, which is all a node's state depends on. */
#ifdef _OPENMP
#pragma omp parallel for collapse(2) schedule(static) if ((zone[0][1] - zone[0][0]) * (zone[1][1] - zone[1][0]) * (zone[2][1] - zone[2][0]) > 4096)
#endif
    for (int i = zone[0][0]; i < zone[0][1]; ++i)
      for (int j = zone[1][0]; j < zone[1][1]; ++j)
        for (int k = zone[2][0]; k < zone[2][1]; ++k) {
          int n1 = idx1(L, i, j, k);
          double sf = lbo_solid_fraction(L, i, j, k, x_loc, y_loc, z_loc, r_lb);
          int id_old = 0;
          double sf_old = 0.0;
          if (tag[p] == L->pid[n1 * 2]) {
            id_old = L->pid[n1 * 2];
            sf_old = L->sf[n1 * 2];
          } else if (tag[p] == L->pid[n1 * 2 + 1]) {
            id_old = L->pid[n1 * 2 + 1];
            sf_old = L->sf[n1 * 2 + 1];
          }
          int decFlag = (sf > 0.00001) + 2 * (sf_old > 0.00001);
          double dx = (double)i - x_loc;
          double dy = (double)j - y_loc;
          double dz = (double)k - z_loc;
          double uNode[3];
          uNode[0] = uTrans[0] + omegaLB[1] * dz - omegaLB[2] * dy;
          uNode[1] = uTrans[1] + omegaLB[2] * dx - omegaLB[0] * dz;
          uNode[2] = uTrans[2] + omegaLB[0] * dy - omegaLB[1] * dx;
          switch (decFlag) {
            case 0: slot_clear(L, n1, tag[p]); break;
            case 1: slot_set(L, n1, tag[p], uNode, sf); break;
            case 2: if (id_old == tag[p]) slot_clear(L, n1, tag[p]); break;
            case 3: if (sf > sf_old || id_old == tag[p]) slot_set(L, n1, tag[p], uNode, sf); break;
          }
        }
  }
}

/* One-rank halo refresh, src/LBM_PSM_MPICOMM.h:79-147 called per axis at
 * src/fix_LBM_PSM.cpp:206-213: the halo layer is copied from the opposite
 * interior layer only on periodic axes; x first, then y, then z, each pass
 * spanning the full halo'd extent of the other axes. */
void lbo_halo_exchange(lbo_lattice* L) {
  const int q = L->q;
  const int nx = L->nx, ny = L->ny, nz = L->nz;
  if (L->periodic[0]) {
    for (int j = 0; j < ny; ++j)
      for (int k = 0; k < nz; ++k) {
        memcpy(L->f_curr + (size_t)idx1(L, 0, j, k) * q, L->f_curr + (size_t)idx1(L, nx - 2, j, k) * q, q * sizeof(double));
        memcpy(L->f_curr + (size_t)idx1(L, nx - 1, j, k) * q, L->f_curr + (size_t)idx1(L, 1, j, k) * q, q * sizeof(double));
      }
  }
  if (L->periodic[1]) {
    for (int i = 0; i < nx; ++i)
      for (int k = 0; k < nz; ++k) {
        memcpy(L->f_curr + (size_t)idx1(L, i, 0, k) * q, L->f_curr + (size_t)idx1(L, i, ny - 2, k) * q, q * sizeof(double));
        memcpy(L->f_curr + (size_t)idx1(L, i, ny - 1, k) * q, L->f_curr + (size_t)idx1(L, i, 1, k) * q, q * sizeof(double));
      }
  }
  if (L->dim == 3 && L->periodic[2]) {
    for (int i = 0; i < nx; ++i)
      for (int j = 0; j < ny; ++j) {
        memcpy(L->f_curr + (size_t)idx1(L, i, j, 0) * q, L->f_curr + (size_t)idx1(L, i, j, nz - 2) * q, q * sizeof(double));
        memcpy(L->f_curr + (size_t)idx1(L, i, j, nz - 1) * q, L->f_curr + (size_t)idx1(L, i, j, 1) * q, q * sizeof(double));
      }
  }
}

/* compute_macro_values, src/LBM_PSM_BGK_dynamics.cpp:48-77 */
static inline void macro(lbo_lattice* L, int n1) {
  const int q = L->q;
  double rho_tmp = 0.0, jx = 0.0, jy = 0.0, jz = 0.0;
  const double* f = L->f_curr + (size_t)n1 * q;
  for (int iq = 0; iq < q; ++iq) {
    double fi = f[iq];
    rho_tmp += fi;
    jx += fi * L->ex[iq];
    jy += fi * L->ey[iq];
    jz += fi * L->ez[iq];
  }
  L->rho[n1] = rho_tmp;
  jx += L->F[0] * 0.5;
  jy += L->F[1] * 0.5;
  jz += L->F[2] * 0.5;
  L->u[n1 * 3] = jx / rho_tmp;
  L->u[n1 * 3 + 1] = jy / rho_tmp;
  L->u[n1 * 3 + 2] = jz / rho_tmp;
}

/* add_Fhyd, src/LBM_PSM_lattice.cpp:355-363 */
static inline void add_Fhyd(lbo_lattice* L, int n1, int pID, double Fhyd, int dir) {
  if (L->pid[n1 * 2] == pID) L->Fh[n1 * 6 + dir] += Fhyd;
  else if (L->pid[n1 * 2 + 1] == pID) L->Fh[n1 * 6 + 3 + dir] += Fhyd;
}

/* collisionAndStream, src/LBM_PSM_BGK_dynamics.h:40-114 (the f0 store at
 * :49 is a write-only array and is left out). */
static inline void collide_one(lbo_lattice* L, int i, int j, int k, int iq, int is, int js, int ks) {
  const int q = L->q;
  const double omega = L->omega;
  int n1 = idx1(L, i, j, k);
  double rho = L->rho[n1];
  double fi = L->f_curr[(size_t)n1 * q + iq];
  double fi_eq = feq_node(L, iq, n1);
  int pID1 = L->pid[n1 * 2];
  int pID2 = L->pid[n1 * 2 + 1];
  double F_lbm_iq = 0.0;
  if (L->Fmag2 > 0.0) F_lbm_iq = (1.0 - 0.5 * omega) * F_iq(L, iq, n1);
  size_t tgt = (size_t)idx1(L, is, js, ks) * q + iq;
  if (pID1 <= 0 && pID2 <= 0) {
    L->f_next[tgt] = fi + (fi_eq - fi) * omega + F_lbm_iq;
    return;
  }
  double B1 = 0.0, B2 = 0.0, solid_coll1 = 0.0, solid_coll2 = 0.0;
  double oneMinusOmega = 1.0 - omega;
  if (pID1 > 0) {
    B1 = L->sf[n1 * 2];
    double f0_solid1 = feq_vel(L, iq, rho, L->us + n1 * 6);
    solid_coll1 = f0_solid1 - fi + oneMinusOmega * (fi - fi_eq);
  }
  if (pID2 > 0) {
    B2 = L->sf[n1 * 2 + 1];
    double f0_solid2 = feq_vel(L, iq, rho, L->us + n1 * 6 + 3);
    solid_coll2 = f0_solid2 - fi + oneMinusOmega * (fi - fi_eq);
  }
  double Btot = B1 + B2;
  if (Btot > 1.0) {
    B1 = B1 / Btot;
    B2 = B2 / Btot;
    Btot = 1.0;
  }
  L->f_next[tgt] = fi + (1.0 - Btot) * (fi_eq - fi) * omega + B1 * solid_coll1 + B2 * solid_coll2 + F_lbm_iq;
  const double ex_i = L->ex[iq];
  const double ey_i = L->ey[iq];
  const double ez_i = L->ez[iq];
  if (pID1 > 0) {
    add_Fhyd(L, n1, pID1, -B1 * solid_coll1 * ex_i, 0);
    add_Fhyd(L, n1, pID1, -B1 * solid_coll1 * ey_i, 1);
    add_Fhyd(L, n1, pID1, -B1 * solid_coll1 * ez_i, 2);
  }
  if (pID2 > 0) {
    add_Fhyd(L, n1, pID2, -B2 * solid_coll2 * ex_i, 0);
    add_Fhyd(L, n1, pID2, -B2 * solid_coll2 * ey_i, 1);
    add_Fhyd(L, n1, pID2, -B2 * solid_coll2 * ez_i, 2);
  }
}

/* macroCollideStream, src/LBM_PSM_BGK_dynamics.cpp:81-109.  Nodes are
 * independent (each (target,iq) has one source), so the outer loop is split
 * over threads without changing any result. */
void lbo_collide_stream(lbo_lattice* L) {
  const int nx = L->nx, ny = L->ny, nz = L->nz, q = L->q;
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j)
      for (int k = 0; k < nz; ++k) {
        macro(L, idx1(L, i, j, k));
        for (int iq = 0; iq < q; ++iq) {
          int is = i + L->ex[iq];
          if (is < 0 || is > nx - 1) continue;
          int js = j + L->ey[iq];
          if (js < 0 || js > ny - 1) continue;
          int ks = k + L->ez[iq];
          if (ks < 0 || ks > nz - 1) continue;
          collide_one(L, i, j, k, iq, is, js, ks);
        }
      }
  double* t = L->f_curr;
  L->f_curr = L->f_next;
  L->f_next = t;
}

/* ---- Zou-He faces, src/LBM_PSM_zou_he_BC.cpp ---------------------------- */
#define GF(o) (f[(o)])
static const double oneOverThree = 1.0 / 3.0;
static const double twoOverThree = 2.0 / 3.0;
static const double oneOverSix = 1.0 / 6.0;

/* :25-40 */
static void zh2_vel_xn(lbo_lattice* L, int ix, int j0, int j1, double ux) {
  for (int j = j0; j <= j1; ++j) {
    double* f = L->f_curr + (size_t)idx1(L, ix, j, 0) * L->q;
    double rho_tmp = 1.0 / (1.0 - ux) * (GF(0) + GF(2) + GF(4) + 2.0 * (GF(3) + GF(6) + GF(7)));
    f[5] = GF(7) + 0.5 * (GF(4) - GF(2)) + rho_tmp * ux * oneOverSix;
    f[8] = GF(6) + 0.5 * (GF(2) - GF(4)) + rho_tmp * ux * oneOverSix;
    f[1] = GF(3) + twoOverThree * rho_tmp * ux;
  }
}
/* :43-58 */
static void zh2_vel_xp(lbo_lattice* L, int ix, int j0, int j1, double ux) {
  for (int j = j0; j <= j1; ++j) {
    double* f = L->f_curr + (size_t)idx1(L, ix, j, 0) * L->q;
    double rho_tmp = 1.0 / (1.0 - ux) * (GF(0) + GF(2) + GF(4) + 2.0 * (GF(1) + GF(5) + GF(8)));
    f[7] = GF(5) + 0.5 * (GF(2) - GF(4)) - rho_tmp * ux * oneOverSix;
    f[6] = GF(8) + 0.5 * (GF(4) - GF(2)) - rho_tmp * ux * oneOverSix;
    f[3] = GF(1) - twoOverThree * rho_tmp * ux;
  }
}
/* :61-76 */
static void zh2_dens_xp(lbo_lattice* L, int ix, int j0, int j1, double rho_bc) {
  for (int j = j0; j <= j1; ++j) {
    double* f = L->f_curr + (size_t)idx1(L, ix, j, 0) * L->q;
    double ux_tmp = -1.0 + 1.0 / rho_bc * (GF(0) + GF(2) + GF(4) + 2.0 * (GF(1) + GF(5) + GF(8)));
    f[6] = GF(8) + 0.5 * (GF(4) - GF(2)) - rho_bc * ux_tmp * oneOverSix;
    f[7] = GF(5) + 0.5 * (GF(2) - GF(4)) - rho_bc * ux_tmp * oneOverSix;
    f[3] = GF(1) - twoOverThree * rho_bc * ux_tmp;
  }
}
/* :79-94 */
static void zh2_vel_yn(lbo_lattice* L, int iy, int i0, int i1, double ux) {
  for (int i = i0; i <= i1; ++i) {
    double* f = L->f_curr + (size_t)idx1(L, i, iy, 0) * L->q;
    double rho_tmp = GF(0) + GF(3) + GF(1) + 2.0 * (GF(8) + GF(4) + GF(7));
    f[6] = GF(8) + 0.5 * (GF(1) - GF(3)) - rho_tmp * ux * 0.5;
    f[5] = GF(7) + 0.5 * (GF(3) - GF(1)) + rho_tmp * ux * 0.5;
    f[2] = GF(4);
  }
}
/* :97-112 */
static void zh2_vel_yp(lbo_lattice* L, int iy, int i0, int i1, double ux) {
  for (int i = i0; i <= i1; ++i) {
    double* f = L->f_curr + (size_t)idx1(L, i, iy, 0) * L->q;
    double rho_tmp = GF(0) + GF(3) + GF(1) + 2.0 * (GF(6) + GF(2) + GF(5));
    f[7] = GF(5) + 0.5 * (GF(1) - GF(3)) - rho_tmp * ux * 0.5;
    f[8] = GF(6) + 0.5 * (GF(3) - GF(1)) + rho_tmp * ux * 0.5;
    f[4] = GF(2);
  }
}
/* :153-185 */
static void zh3_vel_xn(lbo_lattice* L, int ix, int j0, int j1, int k0, int k1, double ux, double uy, double uz) {
  for (int j = j0; j <= j1; ++j)
    for (int k = k0; k <= k1; ++k) {
      double* f = L->f_curr + (size_t)idx1(L, ix, j, k) * L->q;
      double rho_tmp = 1.0 / (1.0 - ux) * (GF(0) + GF(3) + GF(4) + GF(5) + GF(6) + GF(11) + GF(17) + GF(18) + GF(12) +
                                            2.0 * (GF(2) + GF(14) + GF(8) + GF(16) + GF(10)));
      double Nyx = 0.5 * (GF(3) + GF(11) + GF(17) - (GF(4) + GF(18) + GF(12))) - oneOverThree * rho_tmp * uy;
      double Nzx = 0.5 * (GF(5) + GF(18) + GF(11) - (GF(6) + GF(17) + GF(12))) - oneOverThree * rho_tmp * uz;
      f[1] = GF(2) + oneOverThree * rho_tmp * ux;
      f[13] = GF(14) + rho_tmp * oneOverSix * (ux - uy) + Nyx;
      f[7] = GF(8) + rho_tmp * oneOverSix * (ux + uy) - Nyx;
      f[9] = GF(10) + rho_tmp * oneOverSix * (ux + uz) - Nzx;
      f[15] = GF(16) + rho_tmp * oneOverSix * (ux - uz) + Nzx;
    }
}
/* :189-221 */
static void zh3_vel_xp(lbo_lattice* L, int ix, int j0, int j1, int k0, int k1, double ux, double uy, double uz) {
  for (int j = j0; j <= j1; ++j)
    for (int k = k0; k <= k1; ++k) {
      double* f = L->f_curr + (size_t)idx1(L, ix, j, k) * L->q;
      double rho_tmp = 1.0 / (1.0 + ux) * (GF(0) + GF(3) + GF(4) + GF(5) + GF(6) + GF(11) + GF(17) + GF(18) + GF(12) +
                                            2.0 * (GF(1) + GF(7) + GF(13) + GF(9) + GF(15)));
      double Nyx = 0.5 * (GF(3) + GF(11) + GF(17) - (GF(4) + GF(18) + GF(12))) - oneOverThree * rho_tmp * uy;
      double Nzx = 0.5 * (GF(5) + GF(18) + GF(11) - (GF(6) + GF(17) + GF(12))) - oneOverThree * rho_tmp * uz;
      f[2] = GF(1) - oneOverThree * rho_tmp * ux;
      f[14] = GF(13) + rho_tmp * oneOverSix * (-ux + uy) - Nyx;
      f[8] = GF(7) + rho_tmp * oneOverSix * (-ux - uy) + Nyx;
      f[10] = GF(9) + rho_tmp * oneOverSix * (-ux - uz) + Nzx;
      f[16] = GF(15) + rho_tmp * oneOverSix * (-ux + uz) - Nzx;
    }
}
/* :225-257 */
static void zh3_dens_xp(lbo_lattice* L, int ix, int j0, int j1, int k0, int k1, double rho_bc, double uy, double uz) {
  for (int j = j0; j <= j1; ++j)
    for (int k = k0; k <= k1; ++k) {
      double* f = L->f_curr + (size_t)idx1(L, ix, j, k) * L->q;
      double ux = -1.0 + 1.0 / rho_bc * (GF(0) + GF(3) + GF(4) + GF(5) + GF(6) + GF(11) + GF(17) + GF(18) + GF(12) +
                                          2.0 * (GF(1) + GF(7) + GF(13) + GF(9) + GF(15)));
      double Nyx = 0.5 * (GF(3) + GF(11) + GF(17) - (GF(4) + GF(18) + GF(12))) - oneOverThree * rho_bc * uy;
      double Nzx = 0.5 * (GF(5) + GF(18) + GF(11) - (GF(6) + GF(17) + GF(12))) - oneOverThree * rho_bc * uz;
      f[2] = GF(1) - oneOverThree * rho_bc * ux;
      f[14] = GF(13) + rho_bc * oneOverSix * (-ux + uy) - Nyx;
      f[8] = GF(7) + rho_bc * oneOverSix * (-ux - uy) + Nyx;
      f[10] = GF(9) + rho_bc * oneOverSix * (-ux - uz) + Nzx;
      f[16] = GF(15) + rho_bc * oneOverSix * (-ux + uz) - Nzx;
    }
}
/* :261-293 */
static void zh3_vel_yn(lbo_lattice* L, int iy, int i0, int i1, int k0, int k1, double ux, double uy, double uz) {
  for (int i = i0; i <= i1; ++i)
    for (int k = k0; k <= k1; ++k) {
      double* f = L->f_curr + (size_t)idx1(L, i, iy, k) * L->q;
      double rho_tmp = 1.0 / (1.0 - uy) * (GF(0) + GF(1) + GF(2) + GF(5) + GF(6) + GF(9) + GF(15) + GF(16) + GF(10) +
                                            2.0 * (GF(4) + GF(13) + GF(8) + GF(18) + GF(12)));
      double Nxy = 0.5 * (GF(1) + GF(9) + GF(15) - (GF(2) + GF(16) + GF(10))) - oneOverThree * rho_tmp * ux;
      double Nzy = 0.5 * (GF(5) + GF(9) + GF(16) - (GF(6) + GF(15) + GF(10))) - oneOverThree * rho_tmp * uz;
      f[3] = GF(4) + oneOverThree * rho_tmp * uy;
      f[7] = GF(8) + rho_tmp * oneOverSix * (uy + ux) - Nxy;
      f[14] = GF(13) + rho_tmp * oneOverSix * (uy - ux) + Nxy;
      f[11] = GF(12) + rho_tmp * oneOverSix * (uy + uz) - Nzy;
      f[17] = GF(18) + rho_tmp * oneOverSix * (uy - uz) + Nzy;
    }
}
/* :296-328 */
static void zh3_vel_yp(lbo_lattice* L, int iy, int i0, int i1, int k0, int k1, double ux, double uy, double uz) {
  for (int i = i0; i <= i1; ++i)
    for (int k = k0; k <= k1; ++k) {
      double* f = L->f_curr + (size_t)idx1(L, i, iy, k) * L->q;
      double rho_tmp = 1.0 / (1.0 + uy) * (GF(0) + GF(1) + GF(2) + GF(5) + GF(6) + GF(9) + GF(15) + GF(16) + GF(10) +
                                            2.0 * (GF(3) + GF(7) + GF(14) + GF(11) + GF(17)));
      double Nxy = 0.5 * (GF(1) + GF(9) + GF(15) - (GF(2) + GF(16) + GF(10))) - oneOverThree * rho_tmp * ux;
      double Nzy = 0.5 * (GF(5) + GF(9) + GF(16) - (GF(6) + GF(15) + GF(10))) - oneOverThree * rho_tmp * uz;
      f[4] = GF(3) - oneOverThree * rho_tmp * uy;
      f[8] = GF(7) + rho_tmp * oneOverSix * (-uy - ux) + Nxy;
      f[13] = GF(14) + rho_tmp * oneOverSix * (-uy + ux) - Nxy;
      f[12] = GF(11) + rho_tmp * oneOverSix * (-uy - uz) + Nzy;
      f[18] = GF(17) + rho_tmp * oneOverSix * (-uy + uz) - Nzy;
    }
}
/* :331-363 */
static void zh3_vel_zn(lbo_lattice* L, int iz, int i0, int i1, int j0, int j1, double ux, double uy, double uz) {
  for (int i = i0; i <= i1; ++i)
    for (int j = j0; j <= j1; ++j) {
      double* f = L->f_curr + (size_t)idx1(L, i, j, iz) * L->q;
      double rho_tmp = 1.0 / (1.0 - uz) * (GF(0) + GF(1) + GF(2) + GF(3) + GF(4) + GF(7) + GF(13) + GF(14) + GF(8) +
                                            2.0 * (GF(6) + GF(15) + GF(10) + GF(17) + GF(12)));
      double Nxz = 0.5 * (GF(1) + GF(7) + GF(13) - (GF(2) + GF(14) + GF(8))) - oneOverThree * rho_tmp * ux;
      double Nyz = 0.5 * (GF(3) + GF(7) + GF(14) - (GF(4) + GF(13) + GF(8))) - oneOverThree * rho_tmp * uy;
      f[5] = GF(6) + oneOverThree * rho_tmp * uz;
      f[9] = GF(10) + rho_tmp * oneOverSix * (uz + ux) - Nxz;
      f[16] = GF(15) + rho_tmp * oneOverSix * (uz - ux) + Nxz;
      f[11] = GF(12) + rho_tmp * oneOverSix * (uz + uy) - Nyz;
      f[18] = GF(17) + rho_tmp * oneOverSix * (uz - uy) + Nyz;
    }
}
/* :366-398 */
static void zh3_vel_zp(lbo_lattice* L, int iz, int i0, int i1, int j0, int j1, double ux, double uy, double uz) {
  for (int i = i0; i <= i1; ++i)
    for (int j = j0; j <= j1; ++j) {
      double* f = L->f_curr + (size_t)idx1(L, i, j, iz) * L->q;
      double rho_tmp = 1.0 / (1.0 + uz) * (GF(0) + GF(1) + GF(2) + GF(3) + GF(4) + GF(7) + GF(14) + GF(8) + GF(13) +
                                            2.0 * (GF(5) + GF(9) + GF(16) + GF(11) + GF(18)));
      double Nxz = 0.5 * (GF(1) + GF(7) + GF(13) - (GF(2) + GF(14) + GF(8))) - oneOverThree * rho_tmp * ux;
      double Nyz = 0.5 * (GF(3) + GF(7) + GF(14) - (GF(4) + GF(13) + GF(8))) - oneOverThree * rho_tmp * uy;
      f[6] = GF(5) - oneOverThree * rho_tmp * uz;
      f[15] = GF(16) + rho_tmp * oneOverSix * (-uz + ux) - Nxz;
      f[10] = GF(9) + rho_tmp * oneOverSix * (-uz - ux) + Nxz;
      f[17] = GF(18) + rho_tmp * oneOverSix * (-uz + uy) - Nyz;
      f[12] = GF(11) + rho_tmp * oneOverSix * (-uz - uy) + Nyz;
    }
}
#undef GF

/* Test hook: one face function above on the nodes `plane` x [a0,a1] x
 * [b0,b1] (reference halo'd indices; the tangential axes in the order the
 * function takes them), with an arbitrary imposed velocity or density, so
 * that every formula can be checked against the Zou-He conditions even where
 * no `lbm-psm-bc` keyword reaches it with u != 0. */
int lbo_zou_he_face(lbo_lattice* L, int kind, int plane, int a0, int a1, int b0, int b1,
                    double ux, double uy, double uz, double rho) {
  switch (kind) {
    case LBO_ZH2_VEL_XN: zh2_vel_xn(L, plane, a0, a1, ux); return 0;
    case LBO_ZH2_VEL_XP: zh2_vel_xp(L, plane, a0, a1, ux); return 0;
    case LBO_ZH2_DENS_XP: zh2_dens_xp(L, plane, a0, a1, rho); return 0;
    case LBO_ZH2_VEL_YN: zh2_vel_yn(L, plane, a0, a1, ux); return 0;
    case LBO_ZH2_VEL_YP: zh2_vel_yp(L, plane, a0, a1, ux); return 0;
    case LBO_ZH3_VEL_XN: zh3_vel_xn(L, plane, a0, a1, b0, b1, ux, uy, uz); return 0;
    case LBO_ZH3_VEL_XP: zh3_vel_xp(L, plane, a0, a1, b0, b1, ux, uy, uz); return 0;
    case LBO_ZH3_DENS_XP: zh3_dens_xp(L, plane, a0, a1, b0, b1, rho, uy, uz); return 0;
    case LBO_ZH3_VEL_YN: zh3_vel_yn(L, plane, a0, a1, b0, b1, ux, uy, uz); return 0;
    case LBO_ZH3_VEL_YP: zh3_vel_yp(L, plane, a0, a1, b0, b1, ux, uy, uz); return 0;
    case LBO_ZH3_VEL_ZN: zh3_vel_zn(L, plane, a0, a1, b0, b1, ux, uy, uz); return 0;
    case LBO_ZH3_VEL_ZP: zh3_vel_zp(L, plane, a0, a1, b0, b1, ux, uy, uz); return 0;
    default: return -1;
  }
}

/* fix_LBM_PSM_BC::post_force, src/fix_LBM_PSM_BC.cpp:151-402, one rank
 * (every myloc test is true).  e = envelopeWidth = 1. */
void lbo_apply_bc(lbo_lattice* L, int typeBC, double U) {
  const int e = 1;
  const int nx = L->nx, ny = L->ny, nz = L->nz;
  const double rho_outlet = 1.0;
  if (typeBC <= 0) return;
  if (L->dim == 2) {
    switch (typeBC) {
      case 1:
        zh2_vel_yn(L, 0 + e, e, nx - 1 - e, -U);
        zh2_vel_yp(L, ny - 1 - e, e, nx - 1 - e, U);
        break;
      case 2:
        zh2_vel_xn(L, e, e, ny - 1 - e, U);
        zh2_dens_xp(L, nx - 1 - e, e, ny - 1 - e, rho_outlet);
        zh2_vel_yn(L, 0 + e, 1 + e, nx - 2 - e, U);
        zh2_vel_yp(L, ny - 1 - e, 1 + e, nx - 2 - e, U);
        break;
      case 3:
        zh2_vel_yn(L, 0 + e, e, nx - 1 - e, 0.0);
        zh2_vel_yp(L, ny - 1 - e, e, nx - 1 - e, 0.0);
        break;
      case 4:
        zh2_vel_xn(L, e, e, ny - 1 - e, U);
        zh2_dens_xp(L, nx - 1 - e, e, ny - 1 - e, rho_outlet);
        zh2_vel_yn(L, 0 + e, e, nx - 1 - e, 0.0);
        zh2_vel_yp(L, ny - 1 - e, e, nx - 1 - e, 0.0);
        break;
      case 5:
        zh2_vel_xn(L, e, e, ny - 1 - e, 0.0);
        zh2_vel_xp(L, nx - 1 - e, e, ny - 1 - e, 0.0);
        zh2_vel_yn(L, 0 + e, e, nx - 1 - e, 0.0);
        zh2_vel_yp(L, ny - 1 - e, e, nx - 1 - e, 0.0);
        break;
      default: break;
    }
    return;
  }
  switch (typeBC) {
    case 1:
      zh3_vel_yn(L, 0 + e, e, nx - 1 - e, e, nz - 1 - e, -U, 0.0, 0.0);
      zh3_vel_yp(L, ny - 1 - e, e, nx - 1 - e, e, nz - 1 - e, U, 0.0, 0.0);
      break;
    case 2:
      zh3_vel_xn(L, 0 + e, e, ny - 1 - e, e, nz - 1 - e, U, 0.0, 0.0);
      zh3_dens_xp(L, nx - 1 - e, e, ny - 1 - e, e, nz - 1 - e, rho_outlet, 0.0, 0.0);
      zh3_vel_yn(L, 0 + e, e, nx - 1 - e, e, nz - 1 - e, U, 0.0, 0.0);
      zh3_vel_yp(L, ny - 1 - e, e, nx - 1 - e, e, nz - 1 - e, U, 0.0, 0.0);
      zh3_vel_zn(L, 0 + e, e, nx - 1 - e, e, ny - 1 - e, U, 0.0, 0.0);
      zh3_vel_zp(L, nz - 1 - e, e, nx - 1 - e, e, ny - 1 - e, U, 0.0, 0.0);
      break;
    case 3:
      zh3_vel_yn(L, 0 + e, e, nx - 1 - e, e, nz - 1 - e, 0.0, 0.0, 0.0);
      zh3_vel_yp(L, ny - 1 - e, e, nx - 1 - e, e, nz - 1 - e, 0.0, 0.0, 0.0);
      break;
    case 4:
      zh3_vel_xn(L, 0 + e, e, ny - 1 - e, e, nz - 1 - e, U, 0.0, 0.0);
      zh3_dens_xp(L, nx - 1 - e, e, ny - 1 - e, e, nz - 1 - e, rho_outlet, 0.0, 0.0);
      zh3_vel_yn(L, 0 + e, e, nx - 1 - e, e, nz - 1 - e, 0.0, 0.0, 0.0);
      zh3_vel_yp(L, ny - 1 - e, e, nx - 1 - e, e, nz - 1 - e, 0.0, 0.0, 0.0);
      zh3_vel_zn(L, 0 + e, e, nx - 1 - e, e, ny - 1 - e, 0.0, 0.0, 0.0);
      zh3_vel_zp(L, nz - 1 - e, e, nx - 1 - e, e, ny - 1 - e, 0.0, 0.0, 0.0);
      break;
    case 5:
      zh3_vel_xn(L, 0 + e, e, ny - 1 - e, e, nz - 1 - e, 0.0, 0.0, 0.0);
      zh3_vel_xp(L, nx - 1 - e, e, ny - 1 - e, e, nz - 1 - e, 0.0, 0.0, 0.0);
      zh3_vel_yn(L, 0 + e, e, nx - 1 - e, e, nz - 1 - e, 0.0, 0.0, 0.0);
      zh3_vel_yp(L, ny - 1 - e, e, nx - 1 - e, e, nz - 1 - e, 0.0, 0.0, 0.0);
      zh3_vel_zn(L, 0 + e, e, nx - 1 - e, e, ny - 1 - e, 0.0, 0.0, 0.0);
      zh3_vel_zp(L, nz - 1 - e, e, nx - 1 - e, e, ny - 1 - e, 0.0, 0.0, 0.0);
      break;
    default: break;
  }
}

/* calculateHydrodynamicInteractions, src/LBM_PSM_exchangeParticleData.cpp:172-244 */
void lbo_hydro(const lbo_lattice* L, const lbo_units* U, const double origin[3], int tag,
               const double xPart[3], double rp, double* fHydro, double* tHydro,
               double* stresslet) {
  const int e = 1;
  double x_loc = (xPart[0] - origin[0]) / U->dx - L->roff[0];
  double y_loc = (xPart[1] - origin[1]) / U->dx - L->roff[1];
  double z_loc = (xPart[2] - origin[2]) / U->dx - L->roff[2];
  double r_lb = rp / U->dx;
  int zone[3][2] = {{(int)(x_loc - r_lb) - 2, (int)(x_loc + r_lb) + 2},
                    {(int)(y_loc - r_lb) - 2, (int)(y_loc + r_lb) + 2},
                    {(int)(z_loc - r_lb) - 2, (int)(z_loc + r_lb) + 2}};
  for (int a = 0; a < 2; ++a) {
    zone[0][a] = clampi(zone[0][a], e, L->nx - e);
    zone[1][a] = clampi(zone[1][a], e, L->ny - e);
    zone[2][a] = maxi(clampi(maxi(zone[2][a], e), -2147483647, L->nz - e), a);
  }
  const double forceFactor = U->forceFactor;
  const double torqueFactor = U->torqueFactor;
  for (int i = zone[0][0]; i < zone[0][1]; ++i)
    for (int j = zone[1][0]; j < zone[1][1]; ++j)
      for (int k = zone[2][0]; k < zone[2][1]; ++k) {
        int n1 = idx1(L, i, j, k);
        double Fhyd[3];
        if (tag == L->pid[n1 * 2]) {
          Fhyd[0] = L->Fh[n1 * 6]; Fhyd[1] = L->Fh[n1 * 6 + 1]; Fhyd[2] = L->Fh[n1 * 6 + 2];
        } else if (tag == L->pid[n1 * 2 + 1]) {
          Fhyd[0] = L->Fh[n1 * 6 + 3]; Fhyd[1] = L->Fh[n1 * 6 + 4]; Fhyd[2] = L->Fh[n1 * 6 + 5];
        } else {
          continue;
        }
        fHydro[0] += Fhyd[0] * forceFactor;
        fHydro[1] += Fhyd[1] * forceFactor;
        fHydro[2] += Fhyd[2] * forceFactor;
        double dx = (double)i - x_loc;
        double dy = (double)j - y_loc;
        double dz = (double)k - z_loc;
        tHydro[0] += (dy * Fhyd[2] - dz * Fhyd[1]) * torqueFactor;
        tHydro[1] += (dz * Fhyd[0] - dx * Fhyd[2]) * torqueFactor;
        tHydro[2] += (dx * Fhyd[1] - dy * Fhyd[0]) * torqueFactor;
        stresslet[0] -= 0.5 * (dx * Fhyd[0] + dx * Fhyd[0]) * torqueFactor;
        stresslet[1] -= 0.5 * (dy * Fhyd[1] + dy * Fhyd[1]) * torqueFactor;
        stresslet[2] -= 0.5 * (dz * Fhyd[2] + dz * Fhyd[2]) * torqueFactor;
        stresslet[3] -= 0.5 * (dy * Fhyd[0] + dx * Fhyd[1]) * torqueFactor;
        stresslet[4] -= 0.5 * (dz * Fhyd[0] + dx * Fhyd[2]) * torqueFactor;
        stresslet[5] -= 0.5 * (dz * Fhyd[1] + dy * Fhyd[2]) * torqueFactor;
      }
}

/* fix_LBM_PSM::post_force LBM branch, src/fix_LBM_PSM.cpp:197-265, followed
 * by fix_LBM_PSM_BC::post_force.  F/T/S receive the per-entry reductions of
 * every list entry (locals and ghosts); the caller applies += / = and the
 * ghost reverse sum. */
void lbo_fix_step(lbo_lattice* L, const lbo_units* U, const double origin[3], int n,
                  const int* tag, const double* x, const double* v, const double* omega,
                  const double* radius, int bc_keyword, double* F, double* T, double* S) {
  lbo_set_particles(L, U, origin, n, tag, x, v, omega, radius);
  lbo_halo_exchange(L);
  lbo_collide_stream(L);
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1)
#endif
  for (int p = 0; p < n; ++p) {
    double fh[3] = {0.0, 0.0, 0.0}, th[3] = {0.0, 0.0, 0.0};
    double st[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    lbo_hydro(L, U, origin, tag[p], x + 3 * p, radius[p], fh, th, st);
    if (F) for (int d = 0; d < 3; ++d) F[3 * p + d] = fh[d];
    if (T) for (int d = 0; d < 3; ++d) T[3 * p + d] = th[d];
    if (S) for (int d = 0; d < 6; ++d) S[6 * p + d] = st[d];
  }
  lbo_apply_bc(L, bc_keyword, U->u_lb);
}

/* One MPI rank's sub-domain of a decomposed run: ntotal[a] is the sum of
 * the halo'd local sizes of the ranks below this one along axis a and
 * pcoord[a] its coordinate in the rank grid (src/LBM_PSM_lattice.cpp:69-105,
 * src/LBM_PSM_exchangeParticleData.cpp:25-35).  The halos of such a lattice
 * are filled by the caller, as src/LBM_PSM_MPICOMM.h:79-147 does between
 * ranks; create it with no periodic axis so lbo_halo_exchange leaves them. */
void lbo_set_rank(lbo_lattice* L, const int ntotal[3], const int pcoord[3]) {
  for (int d = 0; d < 3; ++d)
    L->roff[d] = (ntotal[d] - ((double)pcoord[d] * 2.0 * (double)1) - 1);
}

double lbo_mass(const lbo_lattice* L) {
  double m = 0.0;
  int i0 = 1, i1 = L->nx - 1, j0 = 1, j1 = L->ny - 1;
  int k0 = (L->dim == 3) ? 1 : 0, k1 = (L->dim == 3) ? L->nz - 1 : 1;
  for (int i = i0; i < i1; ++i)
    for (int j = j0; j < j1; ++j)
      for (int k = k0; k < k1; ++k)
        for (int iq = 0; iq < L->q; ++iq) m += L->f_curr[(size_t)idx1(L, i, j, k) * L->q + iq];
  return m;
}
