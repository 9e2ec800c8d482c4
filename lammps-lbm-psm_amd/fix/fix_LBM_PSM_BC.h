/* `fix lbm-psm-bc` on one MI355X: selects the Zou-He face set that the
 * lbmpsm library applies after streaming in every LBM step.
 *
 * Drop-in for the reference fix (src/fix_LBM_PSM_BC.h,
 * src/fix_LBM_PSM_BC.cpp): same style name and keywords
 *   fix ID group lbm-psm-bc shear|xFlow|channel|channel-velIn-pressOut|closedBox
 * with u_infty = the lattice velocity of `fix lbm-psm` and an outlet density
 * of 1, as src/fix_LBM_PSM_BC.cpp:156-157 sets them.
 */
#ifdef FIX_CLASS
// clang-format off
FixStyle(lbm-psm-bc,fix_LBM_PSM_BC);
// clang-format on
#else

#ifndef LMP_FIX_LBM_PSM_BC_GPU_H
#define LMP_FIX_LBM_PSM_BC_GPU_H

#include "fix.h"

namespace LAMMPS_NS {

class fix_LBM_PSM;

class fix_LBM_PSM_BC : public Fix {
 public:
  fix_LBM_PSM_BC(class LAMMPS*, int, char**);
  ~fix_LBM_PSM_BC() override = default;
  int setmask() override;
  void init() override;
  void post_force(int) override;

  int typeBC;  // LBMPSM_BC_* keyword id

 private:
  fix_LBM_PSM* fixLBMPSM;
};

}  // namespace LAMMPS_NS

#endif
#endif
