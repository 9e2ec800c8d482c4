/* `fix lbm-psm-vti` host side (see fix_LBM_PSM_writeVTI.h).
 *
 * Keywords, defaults and errors as src/fix_LBM_PSM_writeVTI.cpp:14-48; output
 * at init with zero velocity scale and every N steps in pre_force with
 * Uc/u_lb (:57-98); file layout as src/fix_LBM_PSM_writeVTI.h:82-253 for one
 * rank: the folder is recreated when step 0 is written, points are the
 * interior nodes in x-fastest order, the origin is the physical position of
 * the first interior node, and the appended data blocks are each preceded by
 * their 32-bit byte count.
 */
#include "fix_LBM_PSM_writeVTI.h"

#include <cstring>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <sstream>

#include "comm.h"
#include "domain.h"
#include "error.h"
#include "fix_LBM_PSM.h"
#include "lbmpsm.h"
#include "modify.h"
#include "update.h"

using namespace LAMMPS_NS;
using namespace FixConst;

WriteVTI::WriteVTI(LAMMPS* lmp, int narg, char** arg)
    : Fix(lmp, narg, arg), fileName("flow_field"), binary_(true), useDouble_(false),
      fixLBMPSM(nullptr) {
  const char* illegal = "Illegal fix lbm-psm-vti command";
  if (narg < 5) error->all(FLERR, illegal);
  for (int ifix = 0; ifix < modify->nfix; ifix++)
    if (std::strcmp(modify->fix[ifix]->style, "lbm-psm") == 0)
      fixLBMPSM = static_cast<fix_LBM_PSM*>(modify->fix[ifix]);
  if (!fixLBMPSM) error->all(FLERR, "fix lbm-psm-vti requires a fix lbm-psm defined before it");

  for (int iarg = 3; iarg < narg; iarg += 2) {
    if (iarg + 2 > narg) error->all(FLERR, illegal);
    const int v = std::atoi(arg[iarg + 1]);
    if (std::strcmp(arg[iarg], "every") == 0) {
      if (v <= 0) error->all(FLERR, illegal);
      nevery = v;
    } else if (std::strcmp(arg[iarg], "binary") == 0) {
      if (v < 0 || v > 1) error->all(FLERR, illegal);
      binary_ = v != 0;
    } else if (std::strcmp(arg[iarg], "useDouble") == 0) {
      if (v < 0 || v > 1) error->all(FLERR, illegal);
      useDouble_ = v != 0;
    } else {
      error->all(FLERR, illegal);
    }
  }
  fixLBMPSM->request_macros();
}

int WriteVTI::setmask() {
  int mask = 0;
  mask |= PRE_FORCE;
  return mask;
}

void WriteVTI::init() { write(0, 0.0); }

void WriteVTI::pre_force(int) {
  if (update->ntimestep % nevery) return;
  const lbmpsm_config& cfg = fixLBMPSM->config();
  write((long long)update->ntimestep, cfg.Uc / cfg.u_lb);
}

void WriteVTI::write(long long timestep, double u0) {
  lbmpsm_ctx* ctx = fixLBMPSM->lattice();
  if (!ctx) error->all(FLERR, "fix lbm-psm-vti: fix lbm-psm has no lattice");
  lbmpsm_info info;
  lbmpsm_get_info(ctx, &info);
  const size_t nn = (size_t)info.nx * info.ny * info.nz;
  std::vector<double> sf(2 * nn), rho(nn), u(3 * nn);
  if (lbmpsm_get_slots(ctx, nullptr, sf.data(), nullptr, nullptr) != LBMPSM_OK)
    error->all(FLERR, std::string("fix lbm-psm-vti: ") + lbmpsm_last_error(ctx));
  if (lbmpsm_get_macros(ctx, rho.data(), u.data()) != LBMPSM_OK) {
    // no LBM step yet: the equilibrium start, rho = 1 and u = 0
    std::fill(rho.begin(), rho.end(), 1.0);
    std::fill(u.begin(), u.end(), 0.0);
  }
  std::vector<double> B(nn);
  for (size_t n = 0; n < nn; ++n) B[n] = sf[2 * n] + sf[2 * n + 1];
  if (useDouble_) write_typed<double>(timestep, B, rho, u, fixLBMPSM->get_rho(), u0);
  else write_typed<float>(timestep, B, rho, u, fixLBMPSM->get_rho(), u0);
}

template <typename T>
void WriteVTI::write_typed(long long timestep, const std::vector<double>& B,
                           const std::vector<double>& rho, const std::vector<double>& u,
                           double rho0, double u0) {
  namespace fs = std::filesystem;
  const lbmpsm_config& cfg = fixLBMPSM->config();
  lbmpsm_info info;
  lbmpsm_get_info(fixLBMPSM->lattice(), &info);
  const int nx = info.nx, ny = info.ny, nz = info.nz;
  const bool three_d = domain->dimension == 3;
  const std::string typeStr = sizeof(T) == 8 ? "Float64" : "Float32";

  const std::string folder = "flow_field_output_data";
  if (comm->me == 0 && timestep == 0) {
    if (fs::exists(folder)) fs::remove_all(folder);
    fs::create_directory(folder);
  }

  std::ostringstream oss;
  oss << fileName << "." << std::setw(8) << std::setfill('0') << timestep;
  const std::string baseFileName = oss.str();
  const std::string baseTimePath = folder + "/" + baseFileName;

  // points x fastest; library fields are [x][y][z] with z fastest
  const size_t nP = (size_t)nx * ny * nz;
  std::vector<T> B_o(nP), rho_o(nP), u_o(3 * nP);
  size_t id = 0;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i, ++id) {
        const size_t n = ((size_t)i * ny + j) * nz + k;
        const double Bv = B[n];
        B_o[id] = static_cast<T>((Bv > 1.0 ? 1.0 : Bv) * 1.0);
        rho_o[id] = static_cast<T>(rho[n] * rho0);
        u_o[3 * id + 0] = static_cast<T>(u[3 * n + 0] * u0);
        u_o[3 * id + 1] = static_cast<T>(u[3 * n + 1] * u0);
        u_o[3 * id + 2] = three_d ? static_cast<T>(u[3 * n + 2] * u0) : static_cast<T>(0.0);
      }

  // physical position of the first interior node, as the reference's node
  // coordinates give it: origin + i dx - dx (one halo node) with i = 1
  const double dx = cfg.dx;
  const double originX = cfg.origin[0] + 1.0 * dx - dx * 1.0 + 0.0 * dx;
  const double originY = cfg.origin[1] + 1.0 * dx - dx * 1.0 + 0.0 * dx;
  const double originZ = three_d ? cfg.origin[2] + 1.0 * dx - dx * 1.0 + 0.0 * dx : 0.0;
  const int xE = nx - 1, yE = ny - 1, zE = three_d ? nz - 1 : 0;
  const int nzTot = three_d ? nz : 1;
  const std::string extent = "0 " + std::to_string(xE) + " 0 " + std::to_string(yE) + " 0 " +
                             std::to_string(zE);

  const std::string vfname = baseTimePath + "_rank" + std::to_string(comm->me) + ".vti";
  if (binary_) {
    std::ofstream vti(vfname, std::ios::out | std::ios::binary);
    const uint32_t b_sz = (uint32_t)(nP * sizeof(T));
    const uint32_t u_sz = (uint32_t)(nP * 3 * sizeof(T));
    vti << "<?xml version=\"1.0\"?>\n<VTKFile type=\"ImageData\" version=\"1.0\" "
           "byte_order=\"LittleEndian\" header_type=\"UInt32\">\n";
    vti << " <ImageData WholeExtent=\"0 " << nx - 1 << " 0 " << ny - 1 << " 0 " << nzTot - 1
        << "\"" << " Origin=\"" << originX << " " << originY << " " << originZ << "\""
        << " Spacing=\"" << dx << " " << dx << " " << dx << "\">\n";
    vti << "  <Piece Extent=\"" << extent << "\">\n";
    vti << "   <PointData>\n";
    vti << "    <DataArray type=\"" << typeStr
        << "\" Name=\"SolidFraction\" format=\"appended\" offset=\"0\"/>\n";
    vti << "    <DataArray type=\"" << typeStr << "\" Name=\"Density\" format=\"appended\" offset=\""
        << (4 + b_sz) << "\"/>\n";
    vti << "    <DataArray type=\"" << typeStr
        << "\" Name=\"Velocity\" NumberOfComponents=\"3\" format=\"appended\" offset=\""
        << (2 * (4 + b_sz)) << "\"/>\n";
    vti << "   </PointData>\n  </Piece>\n </ImageData>\n <AppendedData encoding=\"raw\">_";
    vti.write(reinterpret_cast<const char*>(&b_sz), 4);
    vti.write(reinterpret_cast<const char*>(B_o.data()), b_sz);
    vti.write(reinterpret_cast<const char*>(&b_sz), 4);
    vti.write(reinterpret_cast<const char*>(rho_o.data()), b_sz);
    vti.write(reinterpret_cast<const char*>(&u_sz), 4);
    vti.write(reinterpret_cast<const char*>(u_o.data()), u_sz);
    vti << "\n </AppendedData>\n</VTKFile>";
  } else {
    std::ofstream vti(vfname);
    vti << "<?xml version=\"1.0\"?>\n<VTKFile type=\"ImageData\" version=\"0.1\">\n";
    vti << " <ImageData WholeExtent=\"0 " << nx - 1 << " 0 " << ny - 1 << " 0 " << nzTot - 1
        << "\"" << " Origin=\"" << originX << " " << originY << " " << originZ << "\""
        << " Spacing=\"" << dx << " " << dx << " " << dx << "\">\n";
    vti << "  <Piece Extent=\"" << extent << "\">\n";
    vti << "   <PointData>\n";
    auto ascii = [&](const char* nm, int comp, const std::vector<T>& data) {
      vti << "    <DataArray type=\"" << typeStr << "\" Name=\"" << nm
          << "\" NumberOfComponents=\"" << comp << "\" format=\"ascii\">\n";
      for (size_t i = 0; i < data.size(); i++) vti << data[i] << (i % 10 == 9 ? "\n" : " ");
      vti << "\n    </DataArray>\n";
    };
    ascii("SolidFraction", 1, B_o);
    ascii("Density", 1, rho_o);
    ascii("Velocity", 3, u_o);
    vti << "   </PointData>\n  </Piece>\n </ImageData>\n</VTKFile>";
  }

  if (comm->me == 0) {
    std::ofstream pvti(baseTimePath + ".pvti");
    pvti << "<?xml version=\"1.0\"?>\n<VTKFile type=\"PImageData\" version=\"0.1\">\n";
    pvti << " <PImageData WholeExtent=\"0 " << nx - 1 << " 0 " << ny - 1 << " 0 " << nzTot - 1
         << "\"" << " Origin=\"" << originX << " " << originY << " " << originZ << "\""
         << " Spacing=\"" << dx << " " << dx << " " << dx << "\">\n";
    pvti << "  <PPointData>\n";
    pvti << "   <PDataArray type=\"" << typeStr << "\" Name=\"SolidFraction\"/>\n";
    pvti << "   <PDataArray type=\"" << typeStr << "\" Name=\"Density\"/>\n";
    pvti << "   <PDataArray type=\"" << typeStr << "\" Name=\"Velocity\" NumberOfComponents=\"3\"/>\n";
    pvti << "  </PPointData>\n";
    pvti << "  <Piece Extent=\"" << extent << "\" Source=\"" << baseFileName << "_rank0.vti\"/>\n";
    pvti << " </PImageData>\n</VTKFile>";
  }
}
