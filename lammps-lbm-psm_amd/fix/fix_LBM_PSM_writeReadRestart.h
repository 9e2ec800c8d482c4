/* `fix lbm-psm-restart` on the lbmpsm library: raw population checkpoints in
 * the reference's format, so that checkpoints written by either code read
 * in the other (src/fix_LBM_PSM_writeReadRestart.{h,cpp}).
 *
 *   fix ID group lbm-psm-restart every N write 0|1 read 0|1
 *
 * File "LBM-PSM-restartFile-processor-<rank>.bin": a 32-bit element count,
 * then the populations of the reference's halo'd lattice as raw doubles
 * ([nx+2][ny+2][nz+2][q], lbmpsm_get_populations_ref).  Written in
 * pre_force every N steps; read at init.
 */
#ifdef FIX_CLASS
// clang-format off
FixStyle(lbm-psm-restart,LBMPSMWriteReadRestart);
// clang-format on
#else

#ifndef LMP_FIX_LBM_PSM_RESTART_GPU_H
#define LMP_FIX_LBM_PSM_RESTART_GPU_H

#include <string>

#include "fix.h"

namespace LAMMPS_NS {

class fix_LBM_PSM;

class LBMPSMWriteReadRestart : public Fix {
 public:
  LBMPSMWriteReadRestart(class LAMMPS*, int, char**);
  ~LBMPSMWriteReadRestart() override = default;
  int setmask() override;
  void init() override;
  void pre_force(int) override;

 private:
  int iWrite, iRead;
  fix_LBM_PSM* fixLBMPSM;
  std::string file_name() const;
};

}  // namespace LAMMPS_NS

#endif
#endif
