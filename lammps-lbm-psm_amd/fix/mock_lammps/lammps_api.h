// Minimal stand-in for the parts of the LAMMPS C++ API that the lbmpsm
// fixes use (Fix base class, Atom, Comm, Domain, Update, Error, Memory,
// Modify, utils).  It exists so the fixes compile and run in this
// repository, which has no LAMMPS; it is test infrastructure, not LAMMPS.
// The member names and signatures follow LAMMPS so that the same fix
// sources compile inside a LAMMPS tree, where the real headers are used.
#pragma once

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <stdexcept>
#include <map>
#include <string>
#include <utility>
#include <vector>

#define FLERR __FILE__, __LINE__

namespace LAMMPS_NS {

typedef int tagint;
typedef int64_t bigint;

class LAMMPS;
class Fix;

class Error {
 public:
  [[noreturn]] void all(const char* file, int line, const std::string& msg) {
    throw std::runtime_error(std::string(file) + ":" + std::to_string(line) + ": " + msg);
  }
};

class Memory {
 public:
  // 2D arrays as LAMMPS allocates them: one contiguous data block plus a
  // row-pointer array, contents kept on growth
  template <typename T>
  T** grow(T**& array, int n1, int n2, const char*) {
    const size_t rows = n1 > 0 ? (size_t)n1 : 0, cols = n2 > 0 ? (size_t)n2 : 0;
    T* old = nullptr;
    size_t oldrows = 0;
    if (array) {
      auto it = blocks_.find((void*)array);
      old = (T*)it->second.first;
      oldrows = it->second.second;
      blocks_.erase(it);
    }
    T* data = (T*)std::calloc(rows * cols + 1, sizeof(T));
    if (old) {
      std::memcpy(data, old, sizeof(T) * (oldrows < rows ? oldrows : rows) * cols);
      std::free(old);
    }
    T** ptrs = (T**)std::malloc(sizeof(T*) * (rows + 1));
    ptrs[0] = data;
    for (size_t i = 0; i < rows; ++i) ptrs[i] = data + i * cols;
    std::free(array);
    array = ptrs;
    blocks_[(void*)array] = {(void*)data, rows};
    return array;
  }
  template <typename T>
  T* grow(T*& array, int n, const char*) {
    array = (T*)std::realloc(array, sizeof(T) * (n > 0 ? n : 1));
    return array;
  }
  template <typename T>
  void destroy(T**& array) {
    if (!array) return;
    auto it = blocks_.find((void*)array);
    std::free(it->second.first);
    blocks_.erase(it);
    std::free(array);
    array = nullptr;
  }
  template <typename T>
  void destroy(T*& array) {
    std::free(array);
    array = nullptr;
  }

 private:
  std::map<void*, std::pair<void*, size_t>> blocks_;
};

class Atom {
 public:
  enum { GROW = 0 };
  int nlocal = 0, nghost = 0, nmax = 0;
  tagint* tag = nullptr;
  double** x = nullptr;
  double** v = nullptr;
  double** f = nullptr;
  double** omega = nullptr;
  double** torque = nullptr;
  double* radius = nullptr;
  // Atom::add_callback: the embedding program registers the fix that asked
  bool callback_requested = false;
  void add_callback(int) { callback_requested = true; }
};

class Domain {
 public:
  int dimension = 3;
  int xperiodic = 1, yperiodic = 1, zperiodic = 1;
  double xprd = 1, yprd = 1, zprd = 1;
  double boxlo[3] = {0, 0, 0};
};

class Comm {
 public:
  int me = 0, nprocs = 1;
  int procgrid[3] = {1, 1, 1};
  int myloc[3] = {0, 0, 0};
  int procneigh[3][2] = {{0, 0}, {0, 0}, {0, 0}};
  // forward / reverse communication of ghost atoms, supplied by the
  // embedding program
  std::function<void()> forward, reverse;
  void forward_comm(void* = nullptr) {
    if (forward) forward();
  }
  void reverse_comm(void* = nullptr) {
    if (reverse) reverse();
  }
};

class Update {
 public:
  bigint ntimestep = 0;
  double dt = 1.0;
};

class Modify {
 public:
  int nfix = 0;
  Fix** fix = nullptr;
  std::vector<Fix*> list;
  void add(Fix* f) {
    list.push_back(f);
    fix = list.data();
    nfix = (int)list.size();
  }
};

class LAMMPS {
 public:
  Memory* memory = new Memory;
  Error* error = new Error;
  Atom* atom = new Atom;
  Update* update = new Update;
  Domain* domain = new Domain;
  Comm* comm = new Comm;
  Modify* modify = new Modify;
  std::string log;
  ~LAMMPS() {
    delete memory;
    delete error;
    delete atom;
    delete update;
    delete domain;
    delete comm;
    delete modify;
  }
};

class Pointers {
 public:
  explicit Pointers(LAMMPS* ptr)
      : lmp(ptr), memory(ptr->memory), error(ptr->error), atom(ptr->atom),
        update(ptr->update), domain(ptr->domain), comm(ptr->comm), modify(ptr->modify) {}
  virtual ~Pointers() = default;

 protected:
  LAMMPS* lmp;
  Memory*& memory;
  Error*& error;
  Atom*& atom;
  Update*& update;
  Domain*& domain;
  Comm*& comm;
  Modify*& modify;
};

namespace FixConst {
enum { POST_FORCE = 1 << 8, PRE_FORCE = 1 << 7 };
}

class Fix : protected Pointers {
 public:
  char* id;
  char* style;
  int igroup = 0, groupbit = 1;
  int nevery = 1;
  int virial_global_flag = 0, virial_peratom_flag = 0, thermo_virial = 0;
  int comm_reverse = 0;
  int vflag_global = 0, vflag_atom = 0;
  double virial[6] = {0, 0, 0, 0, 0, 0};
  std::vector<std::vector<double>> vatom_store;

  Fix(LAMMPS* lmp, int /*narg*/, char** arg) : Pointers(lmp) {
    id = arg[0];
    style = arg[2];
  }
  ~Fix() override = default;
  virtual int setmask() = 0;
  virtual void init() {}
  virtual void setup(int) {}
  virtual void post_force(int) {}
  virtual void pre_force(int) {}
  virtual void grow_arrays(int) {}
  virtual void copy_arrays(int, int, int) {}
  virtual int pack_exchange(int, double*) { return 0; }
  virtual int unpack_exchange(int, double*) { return 0; }
  virtual int pack_reverse_comm(int, int, double*) { return 0; }
  virtual void unpack_reverse_comm(int, int*, double*) {}

  // virial accumulation as Fix::v_init / Fix::v_tally(int, double*)
  void v_init(int vflag) {
    vflag_global = vflag_atom = vflag ? 1 : 0;
    for (double& c : virial) c = 0.0;
    vatom_store.assign(atom->nmax, std::vector<double>(6, 0.0));
  }
  void v_tally(int i, double* v) {
    if (vflag_global)
      for (int c = 0; c < 6; ++c) virial[c] += v[c];
    if (vflag_atom && i < (int)vatom_store.size())
      for (int c = 0; c < 6; ++c) vatom_store[i][c] += v[c];
  }
};

namespace utils {
inline double numeric(const char* file, int line, const char* str, bool, LAMMPS* lmp) {
  char* end = nullptr;
  double v = std::strtod(str, &end);
  if (!end || *end != '\0') lmp->error->all(file, line, std::string("Expected floating point parameter instead of ") + str);
  return v;
}
inline void logmesg(LAMMPS* lmp, const std::string& s) { lmp->log += s; }
}  // namespace utils

}  // namespace LAMMPS_NS
