/*
 * CPU oracle for the LBM-PSM step of `fix lbm-psm` / `fix lbm-psm-bc`.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithm (timnajuch/LAMMPS-LBM-PSM, src/), used as the checker for the HIP
 * library and as the CPU baseline.  Nothing in the product links against it.
 *
 * It keeps the reference's own data layout on purpose: one rank, lattice with
 * a one-node halo on every axis of the active dimension (no halo in z for 2D),
 * array-of-structures populations with the velocity index fastest
 * (src/LBM_PSM_lattice.h:88-90), push streaming with two buffers and a swap
 * (src/LBM_PSM_BGK_dynamics.cpp:81-109), halo refresh by periodic self copy
 * (src/LBM_PSM_MPICOMM.h:79-147) and the two-slot particle record per node
 * (src/LBM_PSM_particleDataOnLattice.h:22-32).
 */
#ifndef LBMPSM_ORACLE_H
#define LBMPSM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Unit conversion, src/LBM_PSM_unitConversion.cpp:15-38. */
typedef struct {
  double rhof, nu, lc, Re, tau;
  int N, dim;
  double Uc_d, Lc_d, cs, csPow2, nu_lb, dx, dx_d, Uc, u_lb, dt_d, tc;
  double forceFactor, torqueFactor, volumeForceFactor;
} lbo_units;

void lbo_units_init(lbo_units* u, double rhof, double nu, double lc, double Re,
                    int N, double tau, int dim);

typedef struct lbo_lattice lbo_lattice;

/* Global node counts (no halo; nz ignored in 2D), periodic flags per axis,
 * BGK tau and the lattice-unit body force.  Returns NULL on bad input. */
lbo_lattice* lbo_create(int dim, int nx, int ny, int nz, const int periodic[3],
                        double tau, const double F_lbm[3]);
void lbo_destroy(lbo_lattice* L);

/* Halo'd local sizes (reference nx/ny/nz after +2 halo). */
void lbo_dims(const lbo_lattice* L, int out[4]); /* nx_h, ny_h, nz_h, q */

/* initialise_dynamics + initialise_domain slot reset. */
void lbo_initialise(lbo_lattice* L, double rho, double ux, double uy, double uz);

/* setParticlesOnLattice for a list of n particles (locals then ghosts). */
void lbo_set_particles(lbo_lattice* L, const lbo_units* U, const double origin[3],
                       int n, const int* tag, const double* x, const double* v,
                       const double* omega, const double* radius);

double lbo_solid_fraction(const lbo_lattice* L, int i, int j, int k, double xP,
                          double yP, double zP, double rP);

void lbo_halo_exchange(lbo_lattice* L);
void lbo_collide_stream(lbo_lattice* L);

/* fix lbm-psm-bc keyword ids: 0 none, 1 shear, 2 xFlow, 3 channel,
 * 4 channel-velIn-pressOut, 5 closedBox. */
void lbo_apply_bc(lbo_lattice* L, int keyword, double u_infty);

/* One Zou-He face function of src/LBM_PSM_zou_he_BC.cpp on the nodes
 * plane x [a0,a1] x [b0,b1] (b ignored in 2D) with any imposed velocity
 * (ux, uy, uz) or density rho (test hook; returns -1 for an unknown kind).
 * 2D velocity faces take ux only; the 3D density face takes rho, uy, uz. */
enum {
  LBO_ZH2_VEL_XN = 0, LBO_ZH2_VEL_XP = 1, LBO_ZH2_DENS_XP = 2, LBO_ZH2_VEL_YN = 3,
  LBO_ZH2_VEL_YP = 4,
  LBO_ZH3_VEL_XN = 10, LBO_ZH3_VEL_XP = 11, LBO_ZH3_DENS_XP = 12, LBO_ZH3_VEL_YN = 13,
  LBO_ZH3_VEL_YP = 14, LBO_ZH3_VEL_ZN = 15, LBO_ZH3_VEL_ZP = 16
};
int lbo_zou_he_face(lbo_lattice* L, int kind, int plane, int a0, int a1, int b0, int b1,
                    double ux, double uy, double uz, double rho);

/* calculateHydrodynamicInteractions for one particle; accumulates into
 * fh[3], th[3], st[6] (the caller zeroes them). */
void lbo_hydro(const lbo_lattice* L, const lbo_units* U, const double origin[3],
               int tag, const double x[3], double rp, double* fh, double* th,
               double* st);

/* One `post_force` LBM step of fix lbm-psm followed by fix lbm-psm-bc:
 * map, halo, collide-stream, per-particle reduction (all n entries), BC. */
void lbo_fix_step(lbo_lattice* L, const lbo_units* U, const double origin[3],
                  int n, const int* tag, const double* x, const double* v,
                  const double* omega, const double* radius, int bc_keyword,
                  double* F, double* T, double* S);

/* Raw views of the halo'd arrays (reference layouts). */
double* lbo_f(lbo_lattice* L);        /* [nx_h*ny_h*nz_h*q] post-stream     */
double* lbo_rho(lbo_lattice* L);      /* [nodes]                            */
double* lbo_u(lbo_lattice* L);        /* [nodes*3]                          */
int*    lbo_slot_tag(lbo_lattice* L); /* [nodes*2]                          */
double* lbo_slot_sf(lbo_lattice* L);  /* [nodes*2]                          */
double* lbo_slot_u(lbo_lattice* L);   /* [nodes*6]                          */
double* lbo_slot_F(lbo_lattice* L);   /* [nodes*6]                          */

/* Sum of all populations over interior nodes (mass). */
/* Make L one rank's sub-domain of a decomposed run (the map and reduction
 * offsets of that rank); its halos are then the caller's to fill. */
void lbo_set_rank(lbo_lattice* L, const int ntotal[3], const int pcoord[3]);

double lbo_mass(const lbo_lattice* L);

/* Number of OpenMP threads the collide loop uses (1 when built without). */
int lbo_threads(void);
/* Threads of the later OpenMP regions (n > 0). */
void lbo_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
