/* `fix lbm-psm` host side over the lbmpsm C ABI (see fix_LBM_PSM.h).
 *
 * Reference behaviour kept, with the reference lines it follows:
 *  - argument parsing and its errors, src/fix_LBM_PSM.cpp:15-83;
 *  - init(): unit conversion, update->dt = phys_time(1)/nevery, the Mach and
 *    cell-width checks, node counts, body force and an equilibrium start at
 *    rho = 1, u = 0 (src/fix_LBM_PSM.cpp:104-179).  Like the reference, every
 *    init() (every `run`) starts the fluid again from equilibrium;
 *  - post_force(): on LBM steps (ntimestep % nevery == 0) forward_comm, map
 *    locals + ghosts, one LBM step, then f/torque += for locals and = for
 *    ghosts, a per-atom store for the steps in between, v_tally of the
 *    stresslet of every atom and reverse_comm; on the other steps the
 *    stored force and torque of each local atom are added
 *    (src/fix_LBM_PSM.cpp:183-264);
 *  - the per-atom array callbacks, src/fix_LBM_PSM.cpp:276-371.
 */
#include "fix_LBM_PSM.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>

#include "atom.h"
#include "comm.h"
#include "domain.h"
#include "error.h"
#include "memory.h"
#include "update.h"

using namespace LAMMPS_NS;
using namespace FixConst;

namespace {
constexpr int kHydroCols = 6;
}

fix_LBM_PSM::fix_LBM_PSM(LAMMPS* lmp, int narg, char** arg)
    : Fix(lmp, narg, arg),
      Nlc(0), lc(0.0), rho(0.0), nu(0.0), Re(0.0), tau(0.7), storage(LBMPSM_STORE_F64),
      device(0), bc_keyword(LBMPSM_BC_NONE), store_macros(0), ctx(nullptr),
      hydrodynamicInteractions(nullptr) {
  const char* illegal = "Illegal fix lbm-psm command";
  if (narg < 15) error->all(FLERR, illegal);

  virial_global_flag = virial_peratom_flag = 1;
  thermo_virial = 1;
  comm_reverse = kHydroCols;
  F_ext[0] = F_ext[1] = F_ext[2] = 0.0;
  std::memset(&cfg, 0, sizeof(cfg));

  // each keyword takes a fixed number of values; a value that is not
  // positive where the reference requires it is an error
  auto need = [&](int iarg, int nvals) {
    if (iarg + nvals + 1 > narg) error->all(FLERR, illegal);
  };
  auto positive = [&](const char* s) {
    double v = utils::numeric(FLERR, s, false, lmp);
    if (v <= 0.0) error->all(FLERR, illegal);
    return v;
  };
  auto positive_int = [&](const char* s) {
    int v = std::atoi(s);
    if (v <= 0) error->all(FLERR, illegal);
    return v;
  };

  for (int iarg = 3; iarg < narg;) {
    const std::string key = arg[iarg];
    if (key == "every") {
      need(iarg, 1);
      nevery = positive_int(arg[iarg + 1]);
      iarg += 2;
    } else if (key == "Nlc") {
      need(iarg, 1);
      Nlc = positive_int(arg[iarg + 1]);
      iarg += 2;
    } else if (key == "lc" || key == "rho" || key == "nu" || key == "Re") {
      need(iarg, 1);
      double v = positive(arg[iarg + 1]);
      (key == "lc" ? lc : key == "rho" ? rho : key == "nu" ? nu : Re) = v;
      iarg += 2;
    } else if (key == "tau") {
      need(iarg, 1);
      tau = utils::numeric(FLERR, arg[iarg + 1], false, lmp);
      if (tau <= 0.5) error->all(FLERR, illegal);
      iarg += 2;
    } else if (key == "Fext") {
      need(iarg, 3);
      for (int d = 0; d < 3; ++d) F_ext[d] = utils::numeric(FLERR, arg[iarg + 1 + d], false, lmp);
      iarg += 4;
    } else if (key == "storage") {  // not in the reference: population storage form
      need(iarg, 1);
      const std::string s = arg[iarg + 1];
      if (s == "f64") storage = LBMPSM_STORE_F64;
      else if (s == "f32") storage = LBMPSM_STORE_F32;
      else if (s == "f16") storage = LBMPSM_STORE_F16;
      else error->all(FLERR, illegal);
      iarg += 2;
    } else if (key == "device") {  // not in the reference: HIP device ordinal
      need(iarg, 1);
      device = std::atoi(arg[iarg + 1]);
      if (device < 0) error->all(FLERR, illegal);
      iarg += 2;
    } else {
      error->all(FLERR, illegal);
    }
  }

  grow_arrays(atom->nmax);
  atom->add_callback(Atom::GROW);
  for (int i = 0; i < atom->nmax; ++i)
    for (int c = 0; c < kHydroCols; ++c) hydrodynamicInteractions[i][c] = 0.0;
}

fix_LBM_PSM::~fix_LBM_PSM() {
  if (ctx) lbmpsm_destroy(ctx);
  memory->destroy(hydrodynamicInteractions);
}

int fix_LBM_PSM::setmask() {
  int mask = 0;
  mask |= POST_FORCE;
  return mask;
}

void fix_LBM_PSM::check(int rc, const char* what) {
  if (rc == LBMPSM_OK) return;
  std::string msg = std::string("fix lbm-psm: ") + what + " failed: " + lbmpsm_last_error(ctx);
  error->all(FLERR, msg);
}

double fix_LBM_PSM::get_rho() { return rho; }

void fix_LBM_PSM::init() {
  // the lattice lives on one GPU; domain decomposition over ranks is not
  // part of this library
  if (comm->nprocs != 1)
    error->all(FLERR, "fix lbm-psm on the GPU runs the whole lattice on one MPI rank");

  const double Fext[3] = {F_ext[0], F_ext[1], F_ext[2]};
  const double boxlo[3] = {domain->boxlo[0], domain->boxlo[1], domain->boxlo[2]};
  const double prd[3] = {domain->xprd, domain->yprd, domain->zprd};
  const int periodic[3] = {domain->xperiodic, domain->yperiodic, domain->zperiodic};
  lbmpsm_config c;
  int rc = lbmpsm_config_from_fix(&c, domain->dimension, Nlc, lc, rho, nu, Re, tau, Fext, boxlo,
                                  prd, periodic);
  if (rc != LBMPSM_OK) error->all(FLERR, lbmpsm_last_error(nullptr));
  c.storage = storage;
  c.device = device;
  cfg = c;

  // phys_time(1) = dt_d * tc, src/LBM_PSM_unitConversion.cpp:60
  update->dt = cfg.dt_d * cfg.tc / (double)nevery;
  if (comm->me == 0) {
    char buf[160];
    std::snprintf(buf, sizeof(buf), "Mach number is Ma = %.2g\n", cfg.u_lb * std::sqrt(3.0));
    utils::logmesg(lmp, buf);
    std::snprintf(buf, sizeof(buf), "DEM timestep is dt = %.6g\n", update->dt);
    utils::logmesg(lmp, buf);
  }

  // a new equilibrium lattice on every init, as the reference builds a new
  // LBMPSMBGKDynamics there (src/fix_LBM_PSM.cpp:164-168)
  if (ctx) {
    lbmpsm_destroy(ctx);
    ctx = nullptr;
  }
  rc = lbmpsm_create(&cfg, &ctx);
  if (rc != LBMPSM_OK) {
    std::string msg = std::string("fix lbm-psm: lattice creation failed: ") +
                      lbmpsm_last_error(nullptr);
    error->all(FLERR, msg);
  }
  // no faces until a `fix lbm-psm-bc` sets them: Modify::init calls its
  // init() after this one (it must be defined after `fix lbm-psm`), and a BC
  // fix that was unfixed sets none, so the new lattice runs without faces,
  // as the reference does once its post_force no longer applies them
  bc_keyword = LBMPSM_BC_NONE;

  for (int i = 0; i < atom->nmax; ++i)
    for (int c2 = 0; c2 < kHydroCols; ++c2) hydrodynamicInteractions[i][c2] = 0.0;
}

void fix_LBM_PSM::set_boundary(int keyword) {
  bc_keyword = keyword;
  if (!ctx) return;  // applied when init() creates the lattice
  check(lbmpsm_set_bc(ctx, keyword, cfg.u_lb, 1.0), "lbmpsm_set_bc");
}

void fix_LBM_PSM::post_force(int vflag) {
  double** f = atom->f;
  double** t = atom->torque;
  const int nlocal = atom->nlocal;

  if (update->ntimestep % nevery) {
    for (int i = 0; i < nlocal; ++i) {
      const double* h = hydrodynamicInteractions[i];
      f[i][0] += h[0];
      f[i][1] += h[1];
      f[i][2] += h[2];
      t[i][0] += h[3];
      t[i][1] += h[4];
      t[i][2] += h[5];
    }
    return;
  }
  if (!ctx) error->all(FLERR, "fix lbm-psm: post_force before init");

  comm->forward_comm();
  const int n = atom->nlocal + atom->nghost;

  const int* tag = nullptr;
  if constexpr (sizeof(tagint) == sizeof(int)) {
    tag = reinterpret_cast<const int*>(atom->tag);
  } else {
    tagbuf.resize(n);
    for (int i = 0; i < n; ++i) {
      if (atom->tag[i] > (tagint)0x7fffffff)
        error->all(FLERR, "fix lbm-psm: atom IDs above 2^31-1 are not supported");
      tagbuf[i] = (int)atom->tag[i];
    }
    tag = tagbuf.data();
  }
  const double* x = n ? atom->x[0] : nullptr;
  const double* v = n ? atom->v[0] : nullptr;
  const double* om = n ? atom->omega[0] : nullptr;
  check(lbmpsm_set_particles(ctx, n, tag, x, v, om, atom->radius), "lbmpsm_set_particles");
  check(lbmpsm_step(ctx, store_macros ? LBMPSM_STEP_STORE_MACROS : 0), "lbmpsm_step");
  Fbuf.resize((size_t)n * 3);
  Tbuf.resize((size_t)n * 3);
  Sbuf.resize((size_t)n * 6);
  check(lbmpsm_get_hydro(ctx, n, Fbuf.data(), Tbuf.data(), Sbuf.data()), "lbmpsm_get_hydro");

  v_init(vflag);
  for (int i = 0; i < n; ++i) {
    const double* fh = &Fbuf[(size_t)i * 3];
    const double* th = &Tbuf[(size_t)i * 3];
    if (i < nlocal) {
      for (int d = 0; d < 3; ++d) {
        f[i][d] += fh[d];
        t[i][d] += th[d];
      }
    } else {  // ghost images: their force reaches the owner in reverse_comm
      for (int d = 0; d < 3; ++d) {
        f[i][d] = fh[d];
        t[i][d] = th[d];
      }
    }
    double* h = hydrodynamicInteractions[i];
    for (int d = 0; d < 3; ++d) {
      h[d] = fh[d];
      h[3 + d] = th[d];
    }
    v_tally(i, &Sbuf[(size_t)i * 6]);
  }
  comm->reverse_comm();
}

void fix_LBM_PSM::grow_arrays(int nmax) {
  memory->grow(hydrodynamicInteractions, nmax, kHydroCols, "fix_LBM_PSM:hydrodynamicInteractions");
}

void fix_LBM_PSM::copy_arrays(int i, int j, int /*delflag*/) {
  for (int c = 0; c < kHydroCols; ++c) hydrodynamicInteractions[j][c] = hydrodynamicInteractions[i][c];
}

int fix_LBM_PSM::pack_exchange(int i, double* buf) {
  for (int c = 0; c < kHydroCols; ++c) buf[c] = hydrodynamicInteractions[i][c];
  return kHydroCols;
}

int fix_LBM_PSM::unpack_exchange(int nlocal, double* buf) {
  for (int c = 0; c < kHydroCols; ++c) hydrodynamicInteractions[nlocal][c] = buf[c];
  return kHydroCols;
}

int fix_LBM_PSM::pack_reverse_comm_size(int n, int /*first*/) { return n * kHydroCols; }

int fix_LBM_PSM::pack_reverse_comm(int n, int first, double* buf) {
  int m = 0;
  for (int i = first; i < first + n; ++i)
    for (int c = 0; c < kHydroCols; ++c) buf[m++] = hydrodynamicInteractions[i][c];
  return m;
}

void fix_LBM_PSM::unpack_reverse_comm(int n, int* list, double* buf) {
  int m = 0;
  for (int i = 0; i < n; ++i)
    for (int c = 0; c < kHydroCols; ++c) hydrodynamicInteractions[list[i]][c] += buf[m++];
}
