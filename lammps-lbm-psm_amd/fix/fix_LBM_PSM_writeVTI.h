/* `fix lbm-psm-vti` on the lbmpsm library: flow-field output in the
 * reference's VTI/PVTI layout (src/fix_LBM_PSM_writeVTI.{h,cpp}).
 *
 *   fix ID group lbm-psm-vti every N [binary 0|1] [useDouble 0|1]
 *
 * Every N steps (pre_force) and once at init, writes
 * flow_field_output_data/flow_field.<step>_rank<r>.vti and the master
 * flow_field.<step>.pvti with three point fields over the lattice nodes:
 * SolidFraction (sum of both slots, clipped at 1), Density (lattice density
 * times the fluid density) and Velocity (lattice velocity times Uc/u_lb;
 * zero in the init file), as Float32 or Float64, raw-appended (binary) or
 * ASCII.  The fields are those of the last LBM step; `fix lbm-psm` keeps its
 * macroscopic fields while this fix exists.
 */
#ifdef FIX_CLASS
// clang-format off
FixStyle(lbm-psm-vti,WriteVTI);
// clang-format on
#else

#ifndef LMP_FIX_LBM_PSM_VTI_GPU_H
#define LMP_FIX_LBM_PSM_VTI_GPU_H

#include <string>
#include <vector>

#include "fix.h"

namespace LAMMPS_NS {

class fix_LBM_PSM;

class WriteVTI : public Fix {
 public:
  WriteVTI(class LAMMPS*, int, char**);
  ~WriteVTI() override = default;
  int setmask() override;
  void init() override;
  void pre_force(int) override;

 private:
  std::string fileName;
  bool binary_, useDouble_;
  fix_LBM_PSM* fixLBMPSM;

  void write(long long timestep, double u0);
  template <typename T>
  void write_typed(long long timestep, const std::vector<double>& B, const std::vector<double>& rho,
                   const std::vector<double>& u, double rho0, double u0);
};

}  // namespace LAMMPS_NS

#endif
#endif
