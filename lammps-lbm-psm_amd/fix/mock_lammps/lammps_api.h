// Minimal stand-in for the parts of the LAMMPS C++ API that the lbmpsm
// fixes use (Fix base class, Atom, Comm, Domain, Update, Error, Memory,
// Modify, utils).  It exists so the fixes compile and run in this
// repository, which has no LAMMPS; it is test infrastructure, not LAMMPS.
// The member names and signatures follow LAMMPS so that the same fix
// sources compile inside a LAMMPS tree, where the real headers are used.
#pragma once

#include <cmath>
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

// single-process stand-in for the MPI calls the styles make
typedef int MPI_Comm;
#define MPI_INT 1
#define MPI_DOUBLE 2
inline int MPI_Bcast(void*, int, int, int, MPI_Comm) { return 0; }

#ifndef MAX
#define MAX(a, b) ((a) > (b) ? (a) : (b))
#endif
#ifndef MIN
#define MIN(a, b) ((a) < (b) ? (a) : (b))
#endif

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
  [[noreturn]] void one(const char* file, int line, const std::string& msg) { all(file, line, msg); }
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
  T** create(T**& array, int n1, int n2, const char* name) {
    array = nullptr;
    return grow(array, n1, n2, name);
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
  int* type = nullptr;
  int ntypes = 1;
  int omega_flag = 1, radius_flag = 1;
  // Atom::add_callback: the embedding program registers the fix that asked
  bool callback_requested = false;
  void add_callback(int) { callback_requested = true; }
};

class Domain {
 public:
  enum { NO_REMAP, X_REMAP, V_REMAP };
  int dimension = 3;
  int xperiodic = 1, yperiodic = 1, zperiodic = 1;
  double xprd = 1, yprd = 1, zprd = 1;
  double boxlo[3] = {0, 0, 0};
};

class Comm {
 public:
  int me = 0, nprocs = 1;
  int ghost_velocity = 1;
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
  void remove(Fix* f) {
    for (size_t i = 0; i < list.size(); ++i)
      if (list[i] == f) {
        list.erase(list.begin() + (long)i);
        break;
      }
    fix = list.data();
    nfix = (int)list.size();
  }
  // fixes whose style matches a pattern; only the "^name" prefix form
  std::vector<Fix*> get_fix_by_style(const std::string& pattern) const;
};

class Force {
 public:
  int newton_pair = 0;
  double vxmu2f = 1.0;  // units lj / si
};

namespace NeighConst {
enum { REQ_DEFAULT = 0, REQ_FULL = 1 << 0 };
}

class NeighList {
 public:
  int inum = 0;
  int* ilist = nullptr;
  int* numneigh = nullptr;
  int** firstneigh = nullptr;
};

class Neighbor {
 public:
  int requested = -1;  // flags of the last request
  void add_request(void*, int flags = NeighConst::REQ_DEFAULT) { requested = flags; }
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
  Force* force = new Force;
  Neighbor* neighbor = new Neighbor;
  MPI_Comm world = 0;
  std::string log;
  ~LAMMPS() {
    delete force;
    delete neighbor;
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
        update(ptr->update), domain(ptr->domain), comm(ptr->comm), modify(ptr->modify),
        force(ptr->force), neighbor(ptr->neighbor), world(ptr->world) {}
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
  Force*& force;
  Neighbor*& neighbor;
  MPI_Comm& world;
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

// Fix styles the lubrication pair looks for (fix deform / fix wall).
class FixDeform : public Fix {
 public:
  int remapflag = Domain::V_REMAP;
  using Fix::Fix;
  int setmask() override { return 0; }
};
class FixWall : public Fix {
 public:
  int xflag = 0;
  using Fix::Fix;
  int setmask() override { return 0; }
};

inline std::vector<Fix*> Modify::get_fix_by_style(const std::string& pattern) const {
  std::vector<Fix*> out;
  const std::string pre = (!pattern.empty() && pattern[0] == '^') ? pattern.substr(1) : pattern;
  for (Fix* f : list)
    if (std::string(f->style).compare(0, pre.size(), pre) == 0) out.push_back(f);
  return out;
}

// Pair base class: the members and helpers a pair style uses.  ev_tally_xyz
// accumulates the global and per-atom virial as LAMMPS does: with
// newton_pair off each local atom of a pair contributes half of the pair's
// virial.
class Pair : protected Pointers {
 public:
  enum { GEOMETRIC, ARITHMETIC, SIXTHPOWER };
  int allocated = 0;
  int** setflag = nullptr;
  double** cutsq = nullptr;
  int single_enable = 1, comm_forward = 0, no_virial_fdotr_compute = 0;
  int offset_flag = 0, mix_flag = GEOMETRIC;
  int evflag = 0, eflag_global = 0, vflag_global = 0, vflag_atom = 0, vflag_fdotr = 0;
  double eng_vdwl = 0.0, eng_coul = 0.0;
  double virial[6] = {0, 0, 0, 0, 0, 0};
  std::vector<std::vector<double>> vatom;
  NeighList* list = nullptr;

  explicit Pair(LAMMPS* lmp) : Pointers(lmp) {}
  ~Pair() override = default;
  virtual void compute(int, int) = 0;
  virtual void settings(int, char**) = 0;
  virtual void coeff(int, char**) = 0;
  virtual double init_one(int, int) = 0;
  virtual void init_style() {}
  virtual void write_restart(FILE*) {}
  virtual void read_restart(FILE*) {}

  double mix_distance(double a, double b) const {
    if (mix_flag == GEOMETRIC) return std::sqrt(a * b);
    if (mix_flag == ARITHMETIC) return 0.5 * (a + b);
    return std::pow(0.5 * (std::pow(a, 6.0) + std::pow(b, 6.0)), 1.0 / 6.0);
  }
  void ev_setup(int eflag, int vflag) {
    evflag = 1;
    eflag_global = eflag ? 1 : 0;
    vflag_global = (vflag & 3) ? 1 : 0;
    vflag_atom = (vflag & 4) ? 1 : 0;
    eng_vdwl = eng_coul = 0.0;
    for (double& c : virial) c = 0.0;
    if (vflag_atom) vatom.assign(atom->nmax, std::vector<double>(6, 0.0));
  }
  void ev_tally_xyz(int i, int j, int nlocal, int newton_pair, double, double, double fx,
                    double fy, double fz, double delx, double dely, double delz) {
    const double v[6] = {delx * fx, dely * fy, delz * fz, delx * fy, delx * fz, dely * fz};
    auto add = [&](double* acc, double w) {
      for (int c = 0; c < 6; ++c) acc[c] += w * v[c];
    };
    if (vflag_global) {
      if (newton_pair) {
        add(virial, 1.0);
      } else {
        if (i < nlocal) add(virial, 0.5);
        if (j < nlocal) add(virial, 0.5);
      }
    }
    if (vflag_atom) {
      if ((newton_pair || i < nlocal) && i < (int)vatom.size()) add(vatom[i].data(), 0.5);
      if ((newton_pair || j < nlocal) && j < (int)vatom.size()) add(vatom[j].data(), 0.5);
    }
  }
};

namespace MathConst {
static constexpr double MY_PI = 3.14159265358979323846;
}

namespace utils {
inline double numeric(const char* file, int line, const char* str, bool, LAMMPS* lmp) {
  char* end = nullptr;
  double v = std::strtod(str, &end);
  if (!end || *end != '\0') lmp->error->all(file, line, std::string("Expected floating point parameter instead of ") + str);
  return v;
}
inline int inumeric(const char* file, int line, const char* str, bool, LAMMPS* lmp) {
  char* end = nullptr;
  long v = std::strtol(str, &end, 10);
  if (!end || *end != '\0') lmp->error->all(file, line, std::string("Expected integer parameter instead of ") + str);
  return (int)v;
}
// "n", "*", "n*", "*n", "m*n" type ranges within [nmin, nmax]
inline void bounds(const char* file, int line, const std::string& str, int nmin, int nmax,
                   int& nlo, int& nhi, Error* error) {
  const size_t star = str.find('*');
  if (star == std::string::npos) {
    nlo = nhi = std::atoi(str.c_str());
  } else {
    nlo = star == 0 ? nmin : std::atoi(str.substr(0, star).c_str());
    nhi = star + 1 == str.size() ? nmax : std::atoi(str.substr(star + 1).c_str());
  }
  if (nlo < nmin || nhi > nmax || nlo > nhi)
    error->all(file, line, "Numeric index " + str + " is out of bounds");
}
inline void sfread(const char* file, int line, void* ptr, size_t size, size_t num, FILE* fp,
                   const char*, Error* error) {
  if (std::fread(ptr, size, num, fp) != num) error->all(file, line, "Unexpected end of restart file");
}
inline void logmesg(LAMMPS* lmp, const std::string& s) { lmp->log += s; }
}  // namespace utils

}  // namespace LAMMPS_NS
