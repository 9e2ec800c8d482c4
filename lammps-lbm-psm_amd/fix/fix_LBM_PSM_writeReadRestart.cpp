/* `fix lbm-psm-restart` host side (see fix_LBM_PSM_writeReadRestart.h).
 * Argument parsing and file format as src/fix_LBM_PSM_writeReadRestart.cpp:
 * 14-41 (keywords), :56-77 (read at init, write in pre_force every N
 * steps), :80-103 (uint32 count, then the doubles).
 */
#include "fix_LBM_PSM_writeReadRestart.h"

#include <cstdint>
#include <cstring>
#include <fstream>
#include <vector>

#include "comm.h"
#include "error.h"
#include "fix_LBM_PSM.h"
#include "lbmpsm.h"
#include "modify.h"
#include "update.h"

using namespace LAMMPS_NS;
using namespace FixConst;

LBMPSMWriteReadRestart::LBMPSMWriteReadRestart(LAMMPS* lmp, int narg, char** arg)
    : Fix(lmp, narg, arg), iWrite(0), iRead(0), fixLBMPSM(nullptr) {
  const char* illegal = "Illegal fix lbm-psm-restart command";
  if (narg < 9) error->all(FLERR, illegal);
  for (int ifix = 0; ifix < modify->nfix; ifix++)
    if (std::strcmp(modify->fix[ifix]->style, "lbm-psm") == 0)
      fixLBMPSM = static_cast<fix_LBM_PSM*>(modify->fix[ifix]);
  if (!fixLBMPSM) error->all(FLERR, "fix lbm-psm-restart requires a fix lbm-psm defined before it");

  for (int iarg = 3; iarg < narg; iarg += 2) {
    if (iarg + 2 > narg) error->all(FLERR, illegal);
    const int v = std::atoi(arg[iarg + 1]);
    if (std::strcmp(arg[iarg], "every") == 0) {
      if (v <= 0) error->all(FLERR, illegal);
      nevery = v;
    } else if (std::strcmp(arg[iarg], "write") == 0) {
      if (v < 0 || v > 1) error->all(FLERR, illegal);
      iWrite = v;
    } else if (std::strcmp(arg[iarg], "read") == 0) {
      if (v < 0 || v > 1) error->all(FLERR, illegal);
      iRead = v;
    } else {
      error->all(FLERR, illegal);
    }
  }
}

int LBMPSMWriteReadRestart::setmask() {
  int mask = 0;
  mask |= PRE_FORCE;
  return mask;
}

std::string LBMPSMWriteReadRestart::file_name() const {
  return "LBM-PSM-restartFile-processor-" + std::to_string(comm->me) + ".bin";
}

void LBMPSMWriteReadRestart::init() {
  if (iRead != 1) return;
  lbmpsm_ctx* ctx = fixLBMPSM->lattice();
  if (!ctx) error->all(FLERR, "fix lbm-psm-restart: fix lbm-psm has no lattice");
  std::ifstream in(file_name(), std::ios::binary);
  if (!in) error->all(FLERR, "fix lbm-psm-restart: cannot open " + file_name());
  uint32_t count = 0;
  in.read(reinterpret_cast<char*>(&count), sizeof(count));
  lbmpsm_info info;
  lbmpsm_get_info(ctx, &info);
  const uint64_t want = (uint64_t)(info.nx + 2) * (info.ny + 2) *
                        (info.dim == 3 ? info.nz + 2 : 1) * info.q;
  if (!in || count != want)
    error->all(FLERR, "fix lbm-psm-restart: " + file_name() + " holds " + std::to_string(count) +
                          " values, the lattice needs " + std::to_string(want));
  std::vector<double> f(count);
  in.read(reinterpret_cast<char*>(f.data()), (std::streamsize)count * sizeof(double));
  if (!in) error->all(FLERR, "fix lbm-psm-restart: " + file_name() + " is truncated");
  if (lbmpsm_set_populations_ref(ctx, f.data()) != LBMPSM_OK)
    error->all(FLERR, std::string("fix lbm-psm-restart: ") + lbmpsm_last_error(ctx));
}

void LBMPSMWriteReadRestart::pre_force(int) {
  if (update->ntimestep % nevery) return;
  if (iWrite != 1) return;
  lbmpsm_ctx* ctx = fixLBMPSM->lattice();
  lbmpsm_info info;
  lbmpsm_get_info(ctx, &info);
  const uint64_t count = (uint64_t)(info.nx + 2) * (info.ny + 2) *
                         (info.dim == 3 ? info.nz + 2 : 1) * info.q;
  if (count > 0xffffffffull) error->all(FLERR, "fix lbm-psm-restart: lattice too large for the format");
  std::vector<double> f(count);
  if (lbmpsm_get_populations_ref(ctx, f.data()) != LBMPSM_OK)
    error->all(FLERR, std::string("fix lbm-psm-restart: ") + lbmpsm_last_error(ctx));
  std::ofstream out(file_name(), std::ios::binary | std::ios::out);
  const uint32_t n32 = (uint32_t)count;
  out.write(reinterpret_cast<const char*>(&n32), sizeof(n32));
  out.write(reinterpret_cast<const char*>(f.data()), (std::streamsize)count * sizeof(double));
}
