// A minimal LAMMPS time-step loop around the lbmpsm fixes, with a C ABI for
// the tests.  It mirrors what LAMMPS does around the force computation for
// an atom_style sphere run with no pair style: fixes are created from their
// command words in deck order, init() runs once per `run`, and every step
// sets the atoms (locals then ghost images, as the test's neighbour build
// hands them over), clears f and torque, calls pre_force and then
// post_force of every fix in order and lets reverse_comm add each ghost's
// force and torque to its owner.  Test infrastructure only.
#include <algorithm>
#include <cstring>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "../fix_LBM_PSM.h"
#include "../fix_LBM_PSM_BC.h"
#include "../fix_LBM_PSM_writeReadRestart.h"
#include "../fix_LBM_PSM_writeVTI.h"
#include "../pair_GRM_LBDEM.h"
#include "lammps_api.h"

using namespace LAMMPS_NS;

struct mock_lmp {
  LAMMPS lmp;
  std::unique_ptr<PairLubricateGRMLBDEM> pair;
  bool pair_ready = false;
  NeighList nlist;
  std::vector<int> ilist, numneigh, neigh;
  std::vector<int*> firstneigh;
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
  mem->grow(a->type, nmax, "atom:type");
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
  mem->destroy(a->type);
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

// `unfix ID`: the fix is deleted and leaves the fix list
int mock_unfix(mock_lmp* m, const char* id) {
  return guarded(m, [&] {
    for (size_t i = 0; i < m->fixes.size(); ++i)
      if (std::strcmp(m->fixes[i]->id, id) == 0) {
        Fix* f = m->fixes[i].get();
        m->lmp.modify->remove(f);
        m->callbacks.erase(std::remove(m->callbacks.begin(), m->callbacks.end(), f),
                           m->callbacks.end());
        m->fixes.erase(m->fixes.begin() + (long)i);
        return;
      }
    throw std::runtime_error(std::string("Could not find fix ID ") + id + " to delete");
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

// ---- pair_style lubricate/GRM/LBDEM -----------------------------------------
namespace {
std::vector<char*> split_words(mock_lmp* m, const char* line) {
  std::istringstream is(line);
  std::vector<std::string> w;
  for (std::string s; is >> s;) w.push_back(s);
  m->words.push_back(w);
  std::vector<char*> argv;
  for (auto& s : m->words.back()) argv.push_back(s.data());
  return argv;
}
}  // namespace

// Atom types, newton_pair, comm_modify vel, units (force->vxmu2f), and the
// deform / wall fixes init_style looks for ("deform v|x", "wall 0|1").
int mock_pair_env(mock_lmp* m, int ntypes, int newton_pair, int ghost_velocity, double vxmu2f,
                  const char* extra_fix) {
  return guarded(m, [&] {
    m->lmp.atom->ntypes = ntypes;
    m->lmp.force->newton_pair = newton_pair;
    m->lmp.comm->ghost_velocity = ghost_velocity;
    m->lmp.force->vxmu2f = vxmu2f;
    if (extra_fix && *extra_fix) {
      auto argv = split_words(m, extra_fix);
      std::vector<char*> fa = {argv[0], argv[0], argv[0]};
      const std::string kind = argv[0];
      if (kind == "deform") {
        auto* f = new FixDeform(&m->lmp, 3, fa.data());
        f->style = (char*)"deform";
        f->remapflag = std::string(argv[1]) == "v" ? Domain::V_REMAP : Domain::X_REMAP;
        m->lmp.modify->add(f);
        m->fixes.emplace_back(f);
      } else if (kind == "wall") {
        auto* f = new FixWall(&m->lmp, 3, fa.data());
        f->style = (char*)"wall/lj93";
        f->xflag = std::atoi(argv[1]);
        m->lmp.modify->add(f);
        m->fixes.emplace_back(f);
      } else {
        throw std::runtime_error("unknown extra fix " + kind);
      }
    }
  });
}

// `pair_style lubricate/GRM/LBDEM args...` without the word "pair_style"
int mock_pair_style(mock_lmp* m, const char* line) {
  return guarded(m, [&] {
    auto argv = split_words(m, line);
    if (argv.empty() || std::string(argv[0]) != "lubricate/GRM/LBDEM")
      throw std::runtime_error("Unrecognized pair style");
    if (!m->pair) m->pair.reset(new PairLubricateGRMLBDEM(&m->lmp));
    m->pair_ready = false;
    m->pair->settings((int)argv.size() - 1, argv.data() + 1);
  });
}

// `pair_coeff I J [cut_inner cut]` without the word "pair_coeff"
int mock_pair_coeff(mock_lmp* m, const char* line) {
  return guarded(m, [&] {
    if (!m->pair) throw std::runtime_error("Pair_coeff command before pair_style is defined");
    auto argv = split_words(m, line);
    m->pair_ready = false;
    m->pair->coeff((int)argv.size(), argv.data());
  });
}

// One force evaluation on a given full neighbour list: atoms (locals then
// ghosts) with their types, neigh holding numneigh[i] partner indices of
// local i in turn.  Pair::init runs first when the settings changed: every
// type pair must be set or mixable, cutsq = cut^2.  vflag 1 tallies the
// global virial.  Returns the locals' f and torque and the virial.
int mock_pair_compute(mock_lmp* m, int nlocal, int nghost, const int* type, const double* x,
                      const double* v, const double* omega, const double* radius,
                      const int* numneigh, const int* neigh, int vflag, double* f_out,
                      double* t_out, double* virial_out) {
  return guarded(m, [&] {
    if (!m->pair) throw std::runtime_error("no pair style");
    const int n = nlocal + nghost;
    grow_atoms(m, n > 0 ? n : 1);
    Atom* a = m->lmp.atom;
    a->nlocal = nlocal;
    a->nghost = nghost;
    if (n) {
      std::memcpy(a->type, type, sizeof(int) * n);
      std::memcpy(a->radius, radius, sizeof(double) * n);
      std::memcpy(a->x[0], x, sizeof(double) * 3 * n);
      std::memcpy(a->v[0], v, sizeof(double) * 3 * n);
      std::memcpy(a->omega[0], omega, sizeof(double) * 3 * n);
      std::memset(a->f[0], 0, sizeof(double) * 3 * n);
      std::memset(a->torque[0], 0, sizeof(double) * 3 * n);
    }
    PairLubricateGRMLBDEM* p = m->pair.get();
    if (!m->pair_ready) {
      p->init_style();
      for (int i = 1; i <= a->ntypes; ++i)
        for (int j = i; j <= a->ntypes; ++j) {
          const double c = p->init_one(i, j);
          p->cutsq[i][j] = p->cutsq[j][i] = c * c;
        }
      m->pair_ready = true;
    }
    int tot = 0;
    for (int i = 0; i < nlocal; ++i) tot += numneigh[i];
    m->ilist.resize(nlocal > 0 ? nlocal : 1);
    m->numneigh.assign(numneigh, numneigh + nlocal);
    m->neigh.assign(neigh, neigh + tot);
    m->firstneigh.resize(n > 0 ? n : 1);
    for (int i = 0, o = 0; i < nlocal; o += numneigh[i], ++i) {
      m->ilist[i] = i;
      m->firstneigh[i] = m->neigh.data() + o;
    }
    m->nlist.inum = nlocal;
    m->nlist.ilist = m->ilist.data();
    m->nlist.numneigh = m->numneigh.data();
    m->nlist.firstneigh = m->firstneigh.data();
    p->list = &m->nlist;
    p->compute(0, vflag);
    if (nlocal) {
      std::memcpy(f_out, a->f[0], sizeof(double) * 3 * nlocal);
      std::memcpy(t_out, a->torque[0], sizeof(double) * 3 * nlocal);
    }
    for (int c = 0; c < 6; ++c) virial_out[c] = p->virial[c];
  });
}

// Neighbour request flags of the last init (NeighConst)
int mock_pair_request(mock_lmp* m) { return m->lmp.neighbor->requested; }

// write_restart / read_restart of the pair's settings and coefficients; the
// read replaces the pair with a fresh instance built from the file
int mock_pair_restart(mock_lmp* m, const char* path, int write) {
  return guarded(m, [&] {
    FILE* fp = std::fopen(path, write ? "wb" : "rb");
    if (!fp) throw std::runtime_error(std::string("cannot open ") + path);
    try {
      if (write) {
        if (!m->pair) throw std::runtime_error("no pair style");
        m->pair->write_restart(fp);
      } else {
        m->pair.reset(new PairLubricateGRMLBDEM(&m->lmp));
        m->pair->read_restart(fp);
        m->pair_ready = false;
      }
    } catch (...) {
      std::fclose(fp);
      throw;
    }
    std::fclose(fp);
  });
}

// pack_comm / unpack_comm round trip of ghost v and omega: packs the atoms
// in list[n] and unpacks them into atoms first..first+n-1
int mock_pair_comm(mock_lmp* m, int n, const int* list, int first, double* buf) {
  return guarded(m, [&] {
    if (!m->pair) throw std::runtime_error("no pair style");
    std::vector<int> l(list, list + n);
    const int per = m->pair->pack_comm(n, l.data(), buf, 0, nullptr);
    if (per != 6) throw std::runtime_error("pack_comm size");
    m->pair->unpack_comm(n, first, buf);
  });
}

// fix adapt: pre_adapt("mu") then adapt(value); returns the index
int mock_pair_adapt(mock_lmp* m, const char* name, double value) {
  if (!m->pair) return -9;
  std::string s(name);
  const int w = m->pair->pre_adapt(s.data(), 1, 1, 1, 1);
  if (w == 0) m->pair->adapt(w, 1, 1, 1, 1, value);
  return w;
}

}  // extern "C\"
