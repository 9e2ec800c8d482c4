/*
 * lbmpsm — C ABI of the MI355X (gfx950) LBM-PSM fluid library behind
 * `fix lbm-psm` and `fix lbm-psm-bc`.
 *
 * One context holds one lattice resident in device memory (populations,
 * per-node particle slots, macroscopic fields) and two HIP streams: the
 * main one, and one on which a new particle list is prepared (upload,
 * geometry, solid fractions) beside the previous step.  All arguments are
 * plain pointers and sizes; host arrays belong to the caller and are only
 * read or written during the call.  Every entry point returns LBMPSM_OK (0)
 * or a negative error code and never throws or exits; lbmpsm_last_error()
 * gives the message.  One context per host thread; the calls' work is
 * ordered as the calls are made, and the lbmpsm_get_* calls synchronise it.
 *
 * Coordinates, velocities, angular velocities, radii, forces, torques and
 * stresslets cross the boundary in the fix's physical units, exactly as
 * LAMMPS holds them (contiguous [n][3] arrays such as atom->x[0]).  The unit
 * conversion factors are those of the reference UnitConversion class
 * (src/LBM_PSM_unitConversion.cpp:15-38); lbmpsm_config_from_fix() fills
 * them from the fix arguments with the same formulas.
 */
#ifndef LBMPSM_H
#define LBMPSM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LBMPSM_OK 0
#define LBMPSM_ERR_ARG (-1)      /* invalid argument                       */
#define LBMPSM_ERR_HIP (-2)      /* HIP runtime error                      */
#define LBMPSM_ERR_STATE (-3)    /* call not valid in the context's state  */
#define LBMPSM_ERR_NOMEM (-4)    /* device or host allocation failed       */
#define LBMPSM_ERR_NODEVICE (-5) /* no usable HIP device                   */
#define LBMPSM_ERR_LIMIT (-6)    /* a documented size limit was exceeded   */

/* Population storage forms. */
#define LBMPSM_STORE_F64 0 /* FP64 populations, FP64 arithmetic (reference form) */
#define LBMPSM_STORE_F32 1 /* FP32 f - w*rho0, FP32 arithmetic                   */
#define LBMPSM_STORE_F16 2 /* FP16 (f - w*rho0)*2^s, FP32 arithmetic             */

/* `fix lbm-psm-bc` keywords, src/fix_LBM_PSM_BC.cpp:24-39. */
#define LBMPSM_BC_NONE 0
#define LBMPSM_BC_SHEAR 1
#define LBMPSM_BC_XFLOW 2
#define LBMPSM_BC_CHANNEL 3
#define LBMPSM_BC_CHANNEL_VELIN_PRESSOUT 4
#define LBMPSM_BC_CLOSEDBOX 5

/* lbmpsm_step() flags. */
#define LBMPSM_STEP_STORE_MACROS 1 /* keep this step's rho and u for lbmpsm_get_macros */

typedef struct {
  int dim;           /* 2 (D2Q9) or 3 (D3Q19)                                 */
  int nx, ny, nz;    /* global node counts without halo; nz ignored for dim 2 */
  int periodic[3];   /* LAMMPS periodicity per axis                           */
  double tau;        /* BGK relaxation time, > 0.5                            */
  double F_lbm[3];   /* body force in lattice units                           */
  double origin[3];  /* domain->boxlo (origin[2] is 0 in 2D)                  */
  double dx;         /* UnitConversion::get_dx                                */
  double Uc, u_lb;   /* velocity conversion  v_lb = v / Uc * u_lb             */
  double tc, dt_d;   /* frequency conversion w_lb = w * tc * dt_d             */
  double force_factor, torque_factor; /* lattice -> physical F and T, S      */
  int storage;       /* LBMPSM_STORE_*                                        */
  int device;        /* HIP device ordinal                                    */
} lbmpsm_config;

typedef struct lbmpsm_ctx lbmpsm_ctx;

/* Lattice geometry as the context stores it. */
typedef struct {
  int dim, q;
  int nx, ny, nz;          /* global interior node counts                   */
  int sx, sy, sz;          /* stored extents (one halo layer kept on each
                              non-periodic axis)                            */
  int ref_offset[3];       /* reference halo'd index = stored index + this  */
  int64_t stored_nodes;
  int storage;
  int64_t device_bytes;    /* device memory held by the context             */
} lbmpsm_info;

/* Fill `cfg` from the `fix lbm-psm` arguments (every Nlc lc rho nu Re tau
 * Fext) and the LAMMPS box, with the reference's formulas:
 * UnitConversion (src/LBM_PSM_unitConversion.cpp:15-38), node counts
 * (src/fix_LBM_PSM.cpp:142-154) and the lattice body force
 * (src/fix_LBM_PSM.cpp:161-162).  Also checks Ma <= 0.3 and that the box is
 * an integer number of cells (src/fix_LBM_PSM.cpp:122-140).
 * Replaces fix_LBM_PSM::init's set-up half, src/fix_LBM_PSM.cpp:104-170. */
int lbmpsm_config_from_fix(lbmpsm_config* cfg, int dim, int Nlc, double lc, double rho,
                           double nu, double Re, double tau, const double Fext[3],
                           const double boxlo[3], const double prd[3],
                           const int periodic[3]);

/* Allocate the lattice on cfg->device.
 * Replaces `new LBMPSMBGKDynamics(...)` + `initialise_domain`,
 * src/fix_LBM_PSM.cpp:164-166. */
int lbmpsm_create(const lbmpsm_config* cfg, lbmpsm_ctx** out);
int lbmpsm_destroy(lbmpsm_ctx* ctx);

/* Message of the last failed call on ctx (or of lbmpsm_create when ctx is
 * NULL).  The pointer stays valid until the next call on that context. */
const char* lbmpsm_last_error(const lbmpsm_ctx* ctx);

int lbmpsm_get_info(const lbmpsm_ctx* ctx, lbmpsm_info* info);

/* Equilibrium start and empty particle slots.
 * Replaces LBMPSMBGKDynamics::initialise_dynamics,
 * src/LBM_PSM_BGK_dynamics.cpp:22-44, with the slot reset of
 * src/LBM_PSM_lattice.cpp:142-186. */
int lbmpsm_init_equilibrium(lbmpsm_ctx* ctx, double rho, double ux, double uy, double uz);

/* Select the `fix lbm-psm-bc` face set applied after every streaming step.
 * u_infty is UnitConversion::get_u_lb(), rho_outlet 1.0 in the reference.
 * Replaces fix_LBM_PSM_BC::post_force, src/fix_LBM_PSM_BC.cpp:151-402. */
int lbmpsm_set_bc(lbmpsm_ctx* ctx, int keyword, double u_infty, double rho_outlet);

/* lbmpsm_set_bc with the face ranges a run of the reference on a
 * procgrid[0] x procgrid[1] x procgrid[2] rank grid applies: every rank
 * applies the faces it holds over the same range of its own local indices
 * (src/fix_LBM_PSM_BC.cpp:223-400), which for the 2D xFlow y faces
 * ([1+e, nx-2-e], :239-242) leaves out the first and last interior node of
 * every rank, not only of the lattice.  No counterpart in the reference (it
 * is the reference's behaviour on that grid): for reproducing results the
 * reference published from multi-rank runs.  {1,1,1} is lbmpsm_set_bc. */
int lbmpsm_set_bc_ranks(lbmpsm_ctx* ctx, int keyword, double u_infty, double rho_outlet,
                        const int procgrid[3]);

/* Map n spheres (LAMMPS order, locals then ghosts) onto the lattice: solid
 * fraction, rigid-body node velocity and the two-slot assignment.
 * Replaces ExchangeParticleData::setParticlesOnLattice,
 * src/LBM_PSM_exchangeParticleData.cpp:24-112. */
int lbmpsm_set_particles(lbmpsm_ctx* ctx, int n, const int* tag, const double* x,
                         const double* v, const double* omega, const double* radius);

/* The same as lbmpsm_set_particles for a list that is already in device
 * memory on the context's device: tag, x, v, omega and radius are device
 * pointers with the same layouts, and max_radius bounds the radii (it sizes
 * the per-entry solid-fraction tables; an entry above it is still mapped
 * exactly, by the slower in-kernel path).  No host transfer and no
 * synchronisation: the arrays are read by device work queued by this call,
 * so they must hold the list when the call is made and stay unchanged until
 * the next lbmpsm_step has been synchronised (any lbmpsm_get_* or
 * lbmpsm_sync).  No counterpart in the reference: for callers whose
 * particle data live on the GPU. */
int lbmpsm_set_particles_device(lbmpsm_ctx* ctx, int n, const int* tag, const double* x,
                                const double* v, const double* omega, const double* radius,
                                double max_radius);

/* One LBM step: collide-stream of every node with PSM weighting and Guo
 * forcing, the selected Zou-He faces, and the per-particle force, torque and
 * stresslet reduction for the list last given to lbmpsm_set_particles.
 * Replaces the LBM branch of fix_LBM_PSM::post_force,
 * src/fix_LBM_PSM.cpp:206-229 (halo refresh, macroCollideStream and the
 * calculateHydrodynamicInteractions loop), and fix_LBM_PSM_BC::post_force. */
int lbmpsm_step(lbmpsm_ctx* ctx, int flags);

/* lbmpsm_step without any allocation, for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, hosts
 * that forbid it).  No counterpart in the reference.  The step is the same
 * (same kernels, same results); the group sums of the force/torque/stresslet
 * reduction go to a static device array of the library (16 MiB per device,
 * held until the process ends) instead of context memory, and the per-entry
 * results go to `out`, caller device memory of [n][12] doubles per list
 * entry: F (3), T (3), S (6: xx yy zz xy xz yz), physical units, n the
 * length of the list last given to lbmpsm_set_particles(_device).
 * Limits and errors: at most 87381 list entries (LBMPSM_ERR_LIMIT beyond);
 * LBMPSM_STEP_STORE_MACROS only after an lbmpsm_step with that flag has
 * allocated the macro arrays (LBMPSM_ERR_STATE otherwise); no profiling.
 * The static array is shared by every call on one device: calls that
 * overlap on one GPU (two contexts on two streams) race.  lbmpsm_get_hydro
 * reports LBMPSM_ERR_STATE after this call until the next lbmpsm_step. */
int lbmpsm_step_noalloc(lbmpsm_ctx* ctx, int flags, double* out);

/* Per-list-entry hydrodynamic force [n][3], torque [n][3] and stresslet
 * [n][6] (xx yy zz xy xz yz) of the last step, physical units; any pointer
 * may be NULL.  Replaces the fh/th/stresslet outputs of
 * ExchangeParticleData::calculateHydrodynamicInteractions,
 * src/LBM_PSM_exchangeParticleData.cpp:172-244. */
int lbmpsm_get_hydro(lbmpsm_ctx* ctx, int n, double* F, double* T, double* S);

/* Populations in the reference's post-stream, post-BC state and layout
 * (src/LBM_PSM_lattice.h:90): interior nodes only, [nx][ny][nz][q] with
 * the velocity index fastest (nz = 1 in 2D).
 * Replaces LBMPSMLattice::getVector_f_curr / setVector_f_curr,
 * src/LBM_PSM_lattice.cpp:250-252. */
int lbmpsm_get_populations(lbmpsm_ctx* ctx, double* f);
int lbmpsm_set_populations(lbmpsm_ctx* ctx, const double* f);

/* The same populations in the reference's full lattice layout, the vector
 * that LBMPSMLattice::getVector_f_curr returns on one rank and that
 * `fix lbm-psm-restart` writes and reads
 * (src/fix_LBM_PSM_writeReadRestart.cpp:80-103): [nx+2][ny+2][nz+2][q]
 * (3D) or [nx+2][ny+2][1][q] (2D) with one halo node on both sides of each
 * axis.  On non-periodic axes the halo is the lattice's own; on periodic
 * axes get writes the periodic image (what the reference's halo refresh
 * puts there before every step) and set ignores it. */
int lbmpsm_get_populations_ref(lbmpsm_ctx* ctx, double* f);
int lbmpsm_set_populations_ref(lbmpsm_ctx* ctx, const double* f);

/* rho [nx][ny][nz] and u [nx][ny][nz][3] of the last step run with
 * LBMPSM_STEP_STORE_MACROS (reference get_rho_reference/get_u_reference,
 * src/LBM_PSM_lattice.cpp:214-222, interior nodes). */
int lbmpsm_get_macros(lbmpsm_ctx* ctx, double* rho, double* u);

/* Per-node particle slots, interior nodes in reference order:
 * tag [node][2], sf [node][2], u_s [node][6], F [node][6] (lattice units)
 * (ParticleDataOnLattice, src/LBM_PSM_particleDataOnLattice.h:22-32);
 * any pointer may be NULL.  A slot set by the last particle list and not
 * yet collided reads F = 0, as the reference's set resets it.  FP32 and
 * FP16 storage keep sf, u_s and F in float on the device; they are returned
 * widened to double. */
int lbmpsm_get_slots(lbmpsm_ctx* ctx, int* tag, double* sf, double* us, double* F);

/* Sum of all populations over interior nodes, accumulated in FP64. */
int lbmpsm_mass(lbmpsm_ctx* ctx, double* mass);

int lbmpsm_sync(lbmpsm_ctx* ctx);

/* Per-kernel timing with HIP events on the context's stream: enable 0 off,
 * 1 every kernel, 2 the bulk collide-stream only (two events per step). */
int lbmpsm_profile_enable(lbmpsm_ctx* ctx, int enable);
/* Total milliseconds and launch count recorded for a kernel name
 * ("collide_stream", "map_particles", "hydro_reduce", ...). */
int lbmpsm_profile_read(lbmpsm_ctx* ctx, const char* kernel, double* total_ms, int64_t* count);
int lbmpsm_profile_reset(lbmpsm_ctx* ctx);

/* ------------------------------------------------------------------------
 * Adjoint (no counterpart in the reference): gradients of a loss on the
 * per-entry force, torque and stresslet after T steps, with the particles
 * frozen (same list, positions and velocities every step), on an FP64,
 * fully periodic context without `lbm-psm-bc` faces.
 *
 * lbmpsm_adjoint_forward runs T = `steps` LBM steps from the context's
 * current state, mapping the list before every step as lbmpsm_set_particles
 * does, and keeps an FP64 copy of the populations every `checkpoint_every`
 * steps.  F, T, S (any may be NULL) receive the last step's per-entry
 * results, as lbmpsm_get_hydro.  Memory: ceil(T/k) + k - 1 population
 * copies.
 *
 * lbmpsm_adjoint_backward takes dLoss/dF [n][3], dLoss/dT [n][3] and
 * dLoss/dS [n][6] of the last step (NULL = zero) and returns
 *   grad_F_lbm[3]   d/d cfg.F_lbm (lattice-unit body force),
 *   grad_tau[1]     d/d cfg.tau (unit conversion factors held fixed),
 *   grad_v[n][3]    d/d v of each list entry (physical units, as passed),
 *   grad_omega[n][3] d/d omega of each list entry.
 * It recomputes each segment from its checkpoint and leaves the populations
 * in the forward's final state.  Call it after lbmpsm_adjoint_forward with
 * no other call on the context in between. */
int lbmpsm_adjoint_forward(lbmpsm_ctx* ctx, int steps, int checkpoint_every, int n,
                           const int* tag, const double* x, const double* v,
                           const double* omega, const double* radius, double* F, double* T,
                           double* S);
int lbmpsm_adjoint_backward(lbmpsm_ctx* ctx, int n, const double* Fbar, const double* Tbar,
                            const double* Sbar, double* grad_F_lbm, double* grad_tau,
                            double* grad_v, double* grad_omega);

#ifdef __cplusplus
}
#endif
#endif
