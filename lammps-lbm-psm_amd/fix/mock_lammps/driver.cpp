// A minimal LAMMPS time-step loop around the lbmpsm fixes, with a C ABI for
// the tests.  It mirrors what LAMMPS does around the force computation for
// an atom_style sphere run with no pair style: fixes are created from their
// command words in deck order, init() runs once per `run`, and every step
// sets the atoms (locals then ghost images, as the test's neighbour build
// hands them over), clears f and torque, calls pre_force and then
// post_force of every fix in order and lets reverse_comm add each ghost's
// force and torque to its owner.  Test infrastructure only.
#include <cstring>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "../fix_LBM_PSM.h"
#include "../fix_LBM_PSM_BC.h"
#include "../fix_LBM_PSM_writeReadRestart.h"
#include "../fix_LBM_PSM_writeVTI.h"
#include "lammps_api.h"

using namespace LAMMPS_NS;

struct mock_lmp {
  LAMMPS lmp;
  std::vector<std::unique_ptr<Fix>> fixes;
  std::vector<Fix*> callbacks;                  // fixes with per-atom arrays
  std::vector<std::vector<std::string>> words;  // keeps argv strings alive
  std::vector<int> owner;                       // ghost -> local index
  std::string err;
};

namespace {

void grow_atoms(mock_lmp* m, int nmax) {
  Atom* a = m->lmp.atom;
  if (nmax <= a->nmax) return;
  Memory* mem = m->lmp.memory;
  mem->grow(a->tag, nmax, "atom:tag");
  mem->grow(a->radius, nmax, "atom:radius");
  mem->grow(a->x, nmax, 3, "atom:x");
  mem->grow(a->v, nmax, 3, "atom:v");
  mem->grow(a->f, nmax, 3, "atom:f");
  mem->grow(a->omega, nmax, 3, "atom:omega");
  mem->grow(a->torque, nmax, 3, "atom:torque");
  a->nmax = nmax;
  for (Fix* f : m->callbacks) f->grow_arrays(nmax);
}

template <typename F>
int guarded(mock_lmp* m, F&& fn) {
  try {
    fn();
    return 0;
  } catch (const std::exception& e) {
    m->err = e.what();
    return -1;
  }
}

}  // namespace

extern "C" {

mock_lmp* mock_open(int dim, const double* boxlo, const double* prd, const int* periodic) {
  auto* m = new mock_lmp;
  Domain* d = m->lmp.domain;
  d->dimension = dim;
  for (int k = 0; k < 3; ++k) d->boxlo[k] = boxlo[k];
  d->xprd = prd[0];
  d->yprd = prd[1];
  d->zprd = prd[2];
  d->xperiodic = periodic[0];
  d->yperiodic = periodic[1];
  d->zperiodic = periodic[2];
  m->lmp.comm->reverse = [m]() {
    Atom* a = m->lmp.atom;
    for (int i = a->nlocal; i < a->nlocal + a->nghost; ++i) {
      const int o = m->owner[i - a->nlocal];
      for (int k = 0; k < 3; ++k) {
        a->f[o][k] += a->f[i][k];
        a->torque[o][k] += a->torque[i][k];
      }
    }
  };
  grow_atoms(m, 1);
  return m;
}

void mock_close(mock_lmp* m) {
  if (!m) return;
  m->fixes.clear();
  Atom* a = m->lmp.atom;
  Memory* mem = m->lmp.memory;
  mem->destroy(a->tag);
  mem->destroy(a->radius);
  mem->destroy(a->x);
  mem->destroy(a->v);
  mem->destroy(a->f);
  mem->destroy(a->omega);
  mem->destroy(a->torque);
  delete m;
}

const char* mock_error(mock_lmp* m) { return m->err.c_str(); }
const char* mock_log(mock_lmp* m) { return m->lmp.log.c_str(); }
double mock_dt(mock_lmp* m) { return m->lmp.update->dt; }
long long mock_ntimestep(mock_lmp* m) { return (long long)m->lmp.update->ntimestep; }

// `fix ID group style args...` without the leading word "fix"
int mock_fix(mock_lmp* m, const char* line) {
  return guarded(m, [&] {
    std::istringstream is(line);
    std::vector<std::string> w;
    for (std::string s; is >> s;) w.push_back(s);
    if (w.size() < 3) throw std::runtime_error("fix command needs ID group style");
    m->words.push_back(w);
    auto& kept = m->words.back();
    std::vector<char*> argv;
    for (auto& s : kept) argv.push_back(s.data());
    const int narg = (int)argv.size();
    m->lmp.atom->callback_requested = false;
    std::unique_ptr<Fix> f;
    if (kept[2] == "lbm-psm")
      f.reset(new fix_LBM_PSM(&m->lmp, narg, argv.data()));
    else if (kept[2] == "lbm-psm-bc")
      f.reset(new fix_LBM_PSM_BC(&m->lmp, narg, argv.data()));
    else if (kept[2] == "lbm-psm-restart")
      f.reset(new LBMPSMWriteReadRestart(&m->lmp, narg, argv.data()));
    else if (kept[2] == "lbm-psm-vti")
      f.reset(new WriteVTI(&m->lmp, narg, argv.data()));
    else
      throw std::runtime_error("Unrecognized fix style '" + kept[2] + "'");
    if (m->lmp.atom->callback_requested) m->callbacks.push_back(f.get());
    m->lmp.modify->add(f.get());
    m->fixes.push_back(std::move(f));
  });
}

// `run`: Modify::init calls init() of every fix in deck order
int mock_init(mock_lmp* m) {
  return guarded(m, [&] {
    for (auto& f : m->fixes) f->init();
  });
}

// One time step.  tag/radius [n], x/v/omega [n][3] with n = nlocal + nghost
// (locals first), owner[nghost] the local index of each ghost.  Returns the
// locals' f and torque [nlocal][3] after post_force and reverse_comm and
// the summed global virial of the fixes.
int mock_step(mock_lmp* m, int nlocal, int nghost, const int* tag, const double* x,
              const double* v, const double* omega, const double* radius, const int* owner,
              double* f_out, double* t_out, double* virial_out) {
  return guarded(m, [&] {
    const int n = nlocal + nghost;
    grow_atoms(m, n > 0 ? n : 1);
    Atom* a = m->lmp.atom;
    a->nlocal = nlocal;
    a->nghost = nghost;
    if (n) {
      std::memcpy(a->tag, tag, sizeof(int) * n);
      std::memcpy(a->radius, radius, sizeof(double) * n);
      std::memcpy(a->x[0], x, sizeof(double) * 3 * n);
      std::memcpy(a->v[0], v, sizeof(double) * 3 * n);
      std::memcpy(a->omega[0], omega, sizeof(double) * 3 * n);
      std::memset(a->f[0], 0, sizeof(double) * 3 * n);
      std::memset(a->torque[0], 0, sizeof(double) * 3 * n);
    }
    m->owner.assign(owner, owner + nghost);
    m->lmp.update->ntimestep += 1;
    for (auto& f : m->fixes)
      if (f->setmask() & FixConst::PRE_FORCE) f->pre_force(1);
    for (auto& f : m->fixes)
      if (f->setmask() & FixConst::POST_FORCE) f->post_force(1);
    if (nlocal) {
      std::memcpy(f_out, a->f[0], sizeof(double) * 3 * nlocal);
      std::memcpy(t_out, a->torque[0], sizeof(double) * 3 * nlocal);
    }
    for (int c = 0; c < 6; ++c) virial_out[c] = 0.0;
    for (auto& f : m->fixes)
      if (f->virial_global_flag)
        for (int c = 0; c < 6; ++c) virial_out[c] += f->virial[c];
  });
}

}  // extern "C"
