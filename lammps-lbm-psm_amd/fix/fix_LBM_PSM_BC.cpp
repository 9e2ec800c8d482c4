/* `fix lbm-psm-bc` host side (see fix_LBM_PSM_BC.h).
 *
 * The reference applies the selected Zou-He faces from its own post_force,
 * which LAMMPS calls right after `fix lbm-psm`'s on the same LBM step
 * (src/fix_LBM_PSM_BC.cpp:151-402).  Here the face set is handed to the
 * library once, at init, and the library applies it inside lbmpsm_step
 * between streaming and the next step, which is the same point of the
 * update: the hydrodynamic reduction that the reference runs in between
 * reads only per-node particle data, never the populations.
 */
#include "fix_LBM_PSM_BC.h"

#include <cstring>
#include <string>

#include "error.h"
#include "fix_LBM_PSM.h"
#include "lbmpsm.h"
#include "modify.h"
#include "update.h"

using namespace LAMMPS_NS;
using namespace FixConst;

fix_LBM_PSM_BC::fix_LBM_PSM_BC(LAMMPS* lmp, int narg, char** arg)
    : Fix(lmp, narg, arg), typeBC(LBMPSM_BC_NONE), fixLBMPSM(nullptr) {
  const char* illegal = "Illegal fix lbm-psm-bc command";
  if (narg < 4) error->all(FLERR, illegal);

  // keyword ids as src/fix_LBM_PSM_BC.cpp:22-39
  static const struct {
    const char* name;
    int id;
  } keywords[] = {{"shear", LBMPSM_BC_SHEAR},
                  {"xFlow", LBMPSM_BC_XFLOW},
                  {"channel", LBMPSM_BC_CHANNEL},
                  {"channel-velIn-pressOut", LBMPSM_BC_CHANNEL_VELIN_PRESSOUT},
                  {"closedBox", LBMPSM_BC_CLOSEDBOX}};
  for (const auto& k : keywords)
    if (std::strcmp(arg[3], k.name) == 0) typeBC = k.id;
  if (typeBC == LBMPSM_BC_NONE) error->all(FLERR, illegal);

  // the most recent `fix lbm-psm` defined before this one
  // (src/fix_LBM_PSM_BC.cpp:124-126)
  for (int ifix = 0; ifix < modify->nfix; ifix++)
    if (std::strcmp(modify->fix[ifix]->style, "lbm-psm") == 0)
      fixLBMPSM = static_cast<fix_LBM_PSM*>(modify->fix[ifix]);
  if (!fixLBMPSM) error->all(FLERR, "fix lbm-psm-bc requires a fix lbm-psm defined before it");
}

int fix_LBM_PSM_BC::setmask() {
  int mask = 0;
  mask |= POST_FORCE;
  return mask;
}

void fix_LBM_PSM_BC::init() {
  // looked up again: the `fix lbm-psm` found at construction may have been
  // unfixed since.  It (defined earlier) has built its lattice in its own
  // init, which leaves the lattice without faces until this call.
  fixLBMPSM = nullptr;
  for (int ifix = 0; ifix < modify->nfix; ifix++)
    if (std::strcmp(modify->fix[ifix]->style, "lbm-psm") == 0)
      fixLBMPSM = static_cast<fix_LBM_PSM*>(modify->fix[ifix]);
  if (!fixLBMPSM) error->all(FLERR, "fix lbm-psm-bc requires a fix lbm-psm defined before it");
  fixLBMPSM->set_boundary(typeBC);
}

void fix_LBM_PSM_BC::post_force(int) {
  // applied inside lbmpsm_step on LBM steps; nothing to do between them
}
