/* `fix lbm-psm` on one MI355X: the LAMMPS-facing host side of the lbmpsm
 * library (include/lbmpsm.h).
 *
 * Drop-in for the reference fix (src/fix_LBM_PSM.h, src/fix_LBM_PSM.cpp):
 * same style name, same arguments, same per-atom force/torque/virial
 * contract, same behaviour between LBM steps.  The lattice, the particle
 * map, collide-stream, the Zou-He faces and the hydrodynamic reduction run
 * in the library on the GPU; this class only moves LAMMPS's atom arrays
 * across the C ABI and folds the results back into atom->f, atom->torque
 * and the fix virial.
 *
 * Syntax (reference syntax plus two optional keywords):
 *   fix ID group lbm-psm every N Nlc n lc L rho R nu NU Re RE
 *       [tau T] [Fext fx fy fz] [storage f64|f32|f16] [device D]
 */
#ifdef FIX_CLASS
// clang-format off
FixStyle(lbm-psm,fix_LBM_PSM);
// clang-format on
#else

#ifndef LMP_FIX_LBM_PSM_GPU_H
#define LMP_FIX_LBM_PSM_GPU_H

#include <vector>

#include "fix.h"
#include "lbmpsm.h"

namespace LAMMPS_NS {

class fix_LBM_PSM : public Fix {
 public:
  fix_LBM_PSM(class LAMMPS*, int, char**);
  ~fix_LBM_PSM() override;
  int setmask() override;
  void init() override;
  void post_force(int) override;

  double get_rho();
  double get_u_lb() const { return cfg.u_lb; }
  double get_Uc() const { return cfg.Uc; }
  double get_dx() const { return cfg.dx; }

  // The library context holding the lattice (valid after init()); the
  // boundary, output and restart fixes act on the lattice through it.
  lbmpsm_ctx* lattice() const { return ctx; }
  const lbmpsm_config& config() const { return cfg; }

  // `fix lbm-psm-bc`: select the Zou-He face set applied inside every LBM
  // step (src/fix_LBM_PSM_BC.cpp:151-402 applies it after collide-stream).
  void set_boundary(int keyword);

  // Keep rho and u of every LBM step (for `fix lbm-psm-vti`).
  void request_macros() { store_macros = 1; }

  void grow_arrays(int) override;
  void copy_arrays(int, int, int) override;
  int pack_exchange(int, double*) override;
  int unpack_exchange(int, double*) override;
  int pack_reverse_comm_size(int, int);
  int pack_reverse_comm(int, int, double*) override;
  void unpack_reverse_comm(int, int*, double*) override;

 private:
  int Nlc;
  double lc, rho, nu, Re, tau;
  double F_ext[3];
  int storage, device;
  int bc_keyword;
  int store_macros;

  lbmpsm_config cfg;
  lbmpsm_ctx* ctx;

  // per-atom force (0-2) and torque (3-5) of the last LBM step, re-applied
  // on the steps in between (src/fix_LBM_PSM.cpp:183-196)
  double** hydrodynamicInteractions;

  // host staging of the list handed to the library
  std::vector<int> tagbuf;
  std::vector<double> Fbuf, Tbuf, Sbuf;

  void check(int rc, const char* what);
};

}  // namespace LAMMPS_NS

#endif
#endif
