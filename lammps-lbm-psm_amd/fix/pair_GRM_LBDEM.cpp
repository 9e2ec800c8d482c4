/* ----------------------------------------------------------------------
   pair_style lubricate/GRM/LBDEM (replaces src/pair_GRM_LBDEM.cpp).

   Host-side DEM pair: the lattice path never calls it.  Forces follow the
   reference term by term, with each resistance expression evaluated in the
   reference's operation order, so that the forces, torques and virial are
   bit-identical to it on the same neighbour list.
------------------------------------------------------------------------- */

#include "pair_GRM_LBDEM.h"

#include <cmath>
#include <cstring>

#include "atom.h"
#include "comm.h"
#include "domain.h"
#include "error.h"
#include "fix_deform.h"
#include "fix_wall.h"
#include "force.h"
#include "math_const.h"
#include "memory.h"
#include "modify.h"
#include "neigh_list.h"
#include "neighbor.h"

using namespace LAMMPS_NS;
using MathConst::MY_PI;

namespace {

// Self and partner coefficient of the selected resistance term at a gap h
// scaled by the mean radius, for beta = a_j / a_i
// (src/pair_GRM_LBDEM.cpp:258-306; Jeffrey & Onishi 1984, leading singular
// terms):
//   1  X^A_11 =  6 pi a_i  2 b^2 / (1+b)^3 / h,            X^A_12 = -X^A_11
//   2  Y^A_11 =  6 pi a_i  4 b (2+b+2b^2) / (15 (1+b)^3) ln(1/h),  Y^A_12 = -Y^A_11
//   3  Y^B_11 = -4 pi a_i^2  b (4+b) / (5 (1+b)^2) ln(1/h),        Y^B_12 = -Y^B_11
//   4  Y^B_11 as in 3, Y^B_21 = -4 pi a_j^2 (1/b)(4+1/b) / (5 (1+1/b)^2) ln(1/h)
//   5  Y^C_11 =  8 pi a_i^3  2b / (5 (1+b)) ln(1/h),  Y^C_12 = 8 pi a_i^3 b^2 / (10 (1+b)) ln(1/h)
struct Resistance {
  double self = 0.0, partner = 0.0;
};

Resistance resistance(int flag, double beta, double ai, double aj, double h) {
  Resistance r;
  const double b2 = beta * beta;
  const double b1 = 1.0 + beta;
  const double b1p2 = b1 * b1;
  const double b1p3 = b1p2 * b1;
  const double lg = log(1.0 / h);
  switch (flag) {
    case 1:
      r.self = 2.0 * b2 / b1p3 / h;
      r.self *= 6.0 * MY_PI * ai;
      r.partner = -r.self;
      break;
    case 2:
      r.self = 4.0 * beta * (2.0 + beta + 2.0 * b2) / (15.0 * b1p3) * lg;
      r.self *= 6.0 * MY_PI * ai;
      r.partner = -r.self;
      break;
    case 3:
      r.self = -beta * (4.0 + beta) / (5.0 * b1p2) * lg;
      r.self *= 4.0 * MY_PI * (ai * ai);
      r.partner = -r.self;
      break;
    case 4: {
      r.self = -beta * (4.0 + beta) / (5.0 * b1p2) * lg;
      r.self *= 4.0 * MY_PI * (ai * ai);
      const double bi = 1.0 / beta;
      const double bi1 = 1.0 + bi;
      r.partner = bi * (4.0 + bi) / (5.0 * (bi1 * bi1)) * lg;
      r.partner *= -4.0 * MY_PI * (aj * aj);
      break;
    }
    case 5: {
      const double ai3 = (ai * ai) * ai;
      r.self = 2.0 * beta / (5.0 * b1) * lg;
      r.self *= 8.0 * MY_PI * ai3;
      r.partner = b2 / (10.0 * b1) * lg;
      r.partner *= 8.0 * MY_PI * ai3;
      break;
    }
    default:
      break;
  }
  return r;
}

inline double dot3(const double *a, const double *e) { return a[0] * e[0] + a[1] * e[1] + a[2] * e[2]; }

}    // namespace

PairLubricateGRMLBDEM::PairLubricateGRMLBDEM(LAMMPS *lmp) : Pair(lmp)
{
  single_enable = 0;
  comm_forward = 6;    // v and omega of ghosts
  no_virial_fdotr_compute = 1;
}

PairLubricateGRMLBDEM::~PairLubricateGRMLBDEM()
{
  if (allocated) {
    memory->destroy(setflag);
    memory->destroy(cutsq);
    memory->destroy(cut);
    memory->destroy(cut_inner);
  }
}

/* ----------------------------------------------------------------------
   src/pair_GRM_LBDEM.cpp:78-476.  Velocities enter with a minus sign
   (U(x) - U_p of a quiescent fluid); e points from i to j.  The correction
   at the gap minus the correction at h_N is added to atom i only (full
   list, newton off) and tallied to i alone.
------------------------------------------------------------------------- */

void PairLubricateGRMLBDEM::compute(int eflag, int vflag)
{
  if (eflag || vflag) ev_setup(eflag, vflag);
  else evflag = vflag_fdotr = 0;

  double **x = atom->x, **v = atom->v, **f = atom->f;
  double **omega = atom->omega, **torque = atom->torque;
  const double *radius = atom->radius;
  const int *type = atom->type;
  const int nlocal = atom->nlocal;
  const double vxmu2f = force->vxmu2f;
  const int flag = flagGRM;

  for (int ii = 0; ii < list->inum; ii++) {
    const int i = list->ilist[ii];
    const int itype = type[i];
    const double ai = radius[i];
    const double vi[3] = {-v[i][0], -v[i][1], -v[i][2]};
    const double wi[3] = {-omega[i][0], -omega[i][1], -omega[i][2]};
    const int *jlist = list->firstneigh[i];
    const int jnum = list->numneigh[i];

    for (int jj = 0; jj < jnum; jj++) {
      const int j = jlist[jj];
      const int jtype = type[j];
      const double delx = x[j][0] - x[i][0];
      const double dely = x[j][1] - x[i][1];
      const double delz = x[j][2] - x[i][2];
      const double rsq = delx * delx + dely * dely + delz * delz;
      const double aj = radius[j];
      const double gapN = sqrt(cutsq[itype][jtype]) * cutoff_factor;
      const double r = sqrt(rsq);
      double gap = r - ai - aj;
      if (!(gap < gapN)) continue;

      const double vj[3] = {-v[j][0], -v[j][1], -v[j][2]};
      const double wj[3] = {-omega[j][0], -omega[j][1], -omega[j][2]};
      // closer than the inner cutoff: the inner cutoff's value
      if (gap < cut_inner[itype][jtype]) gap = cut_inner[itype][jtype];
      const double amean = 0.5 * (ai + aj);
      const double h = gap / amean;
      const double hN = gapN / amean;
      const double beta = aj / ai;
      const Resistance R = resistance(flag, beta, ai, aj, h);
      const Resistance RN = resistance(flag, beta, ai, aj, hN);
      const double e[3] = {delx / r, dely / r, delz / r};

      // one component of the relative motion per term, then
      // mu ((self a_i + partner a_j) - (selfN a_i + partnerN a_j))
      double si[3] = {0.0, 0.0, 0.0}, sj[3] = {0.0, 0.0, 0.0};
      switch (flag) {
        case 1: {    // normal velocities
          const double vin = dot3(vi, e), vjn = dot3(vj, e);
          for (int k = 0; k < 3; k++) {
            si[k] = vin * e[k];
            sj[k] = vjn * e[k];
          }
          break;
        }
        case 2: {    // tangential velocities
          const double vin = dot3(vi, e), vjn = dot3(vj, e);
          for (int k = 0; k < 3; k++) {
            si[k] = vi[k] - vin * e[k];
            sj[k] = vj[k] - vjn * e[k];
          }
          break;
        }
        case 3:    // v x e
          si[0] = vi[1] * e[2] - vi[2] * e[1];
          si[1] = vi[2] * e[0] - vi[0] * e[2];
          si[2] = vi[0] * e[1] - vi[1] * e[0];
          sj[0] = vj[1] * e[2] - vj[2] * e[1];
          sj[1] = vj[2] * e[0] - vj[0] * e[2];
          sj[2] = vj[0] * e[1] - vj[1] * e[0];
          break;
        case 4:    // e x omega
          si[0] = e[1] * wi[2] - e[2] * wi[1];
          si[1] = e[2] * wi[0] - e[0] * wi[2];
          si[2] = e[0] * wi[1] - e[1] * wi[0];
          sj[0] = e[1] * wj[2] - e[2] * wj[1];
          sj[1] = e[2] * wj[0] - e[0] * wj[2];
          sj[2] = e[0] * wj[1] - e[1] * wj[0];
          break;
        case 5: {    // tangential angular velocities
          const double win = dot3(wi, e), wjn = dot3(wj, e);
          for (int k = 0; k < 3; k++) {
            si[k] = wi[k] - win * e[k];
            sj[k] = wj[k] - wjn * e[k];
          }
          break;
        }
        default:
          break;
      }
      double out[3];
      for (int k = 0; k < 3; k++)
        out[k] = mu * ((R.self * si[k] + R.partner * sj[k]) - (RN.self * si[k] + RN.partner * sj[k]));

      double fx = 0.0, fy = 0.0, fz = 0.0;
      if (flag == 1 || flag == 2 || flag == 4) {
        fx = out[0] * vxmu2f;
        fy = out[1] * vxmu2f;
        fz = out[2] * vxmu2f;
        f[i][0] += fx;
        f[i][1] += fy;
        f[i][2] += fz;
      } else if (flag == 3 || flag == 5) {
        torque[i][0] += vxmu2f * out[0];
        torque[i][1] += vxmu2f * out[1];
        torque[i][2] += vxmu2f * out[2];
      }
      // j = nlocal: only i is tallied
      if (evflag) ev_tally_xyz(i, nlocal, nlocal, 0, 0.0, 0.0, fx, fy, fz, -delx, -dely, -delz);
    }
  }
}

void PairLubricateGRMLBDEM::allocate()
{
  allocated = 1;
  const int n = atom->ntypes;
  memory->create(setflag, n + 1, n + 1, "pair:setflag");
  for (int i = 1; i <= n; i++)
    for (int j = i; j <= n; j++) setflag[i][j] = 0;
  memory->create(cutsq, n + 1, n + 1, "pair:cutsq");
  memory->create(cut, n + 1, n + 1, "pair:cut");
  memory->create(cut_inner, n + 1, n + 1, "pair:cut_inner");
}

// pair_style lubricate/GRM/LBDEM mu cut_inner cut flagGRM cutoff_factor
// (src/pair_GRM_LBDEM.cpp:502-522)
void PairLubricateGRMLBDEM::settings(int narg, char **arg)
{
  if (narg != 5) error->all(FLERR, "Illegal pair_style command");
  mu = utils::numeric(FLERR, arg[0], false, lmp);
  cut_inner_global = utils::numeric(FLERR, arg[1], false, lmp);
  cut_global = utils::numeric(FLERR, arg[2], false, lmp);
  flagGRM = utils::inumeric(FLERR, arg[3], false, lmp);
  cutoff_factor = utils::numeric(FLERR, arg[4], false, lmp);
  // type pairs set explicitly before take the new global cutoffs
  if (allocated)
    for (int i = 1; i <= atom->ntypes; i++)
      for (int j = i; j <= atom->ntypes; j++)
        if (setflag[i][j]) {
          cut_inner[i][j] = cut_inner_global;
          cut[i][j] = cut_global;
        }
}

// pair_coeff I J [cut_inner cut]  (src/pair_GRM_LBDEM.cpp:528-556)
void PairLubricateGRMLBDEM::coeff(int narg, char **arg)
{
  if (narg != 2 && narg != 4) error->all(FLERR, "Incorrect args for pair coefficients");
  if (!allocated) allocate();
  int ilo, ihi, jlo, jhi;
  utils::bounds(FLERR, arg[0], 1, atom->ntypes, ilo, ihi, error);
  utils::bounds(FLERR, arg[1], 1, atom->ntypes, jlo, jhi, error);
  const double ci = narg == 4 ? utils::numeric(FLERR, arg[2], false, lmp) : cut_inner_global;
  const double co = narg == 4 ? utils::numeric(FLERR, arg[3], false, lmp) : cut_global;
  int count = 0;
  for (int i = ilo; i <= ihi; i++)
    for (int j = MAX(jlo, i); j <= jhi; j++) {
      cut_inner[i][j] = ci;
      cut[i][j] = co;
      setflag[i][j] = 1;
      count++;
    }
  if (count == 0) error->all(FLERR, "Incorrect args for pair coefficients");
}

// src/pair_GRM_LBDEM.cpp:562-630
void PairLubricateGRMLBDEM::init_style()
{
  if (force->newton_pair == 1) error->all(FLERR, "Pair lubricate/GRM/LBDEM requires newton pair off");
  if (comm->ghost_velocity == 0)
    error->all(FLERR, "Pair lubricate/GRM/LBDEM requires ghost atoms store velocity");
  if (!atom->omega_flag) error->all(FLERR, "Pair lubricate/GRM/LBDEM requires atom attribute omega");
  if (!atom->radius_flag) error->all(FLERR, "Pair lubricate/GRM/LBEM requires atom attribute radius");
  for (int i = 0; i < atom->nlocal; i++)
    if (atom->radius[i] == 0.0) error->one(FLERR, "Pair lubricate/GRM/LBDEM requires extended particles");

  neighbor->add_request(this, NeighConst::REQ_FULL);

  // fix deform must remap velocities; at most one fix wall
  shearing = flagdeform = flagwall = 0;
  auto deforms = modify->get_fix_by_style("^deform");
  if (!deforms.empty()) {
    shearing = flagdeform = 1;
    auto *d = dynamic_cast<FixDeform *>(deforms[0]);
    if (d && d->remapflag != Domain::V_REMAP)
      error->all(FLERR, "Using pair lubricate/GRM/LBDEM with inconsistent fix deform remap option");
  }
  auto walls = modify->get_fix_by_style("^wall");
  if (walls.size() > 1)
    error->all(FLERR, "Cannot use multiple fix wall commands with pair lubricate/GRM/LBDEM");
  if (walls.size() == 1) {
    auto *w = dynamic_cast<FixWall *>(walls[0]);
    if (!w)
      error->all(FLERR, std::string("Fix ") + walls[0]->style +
                            " is not compatible with pair lubricate/GRM/LBDEM");
    flagwall = w->xflag ? 2 : 1;
  }
}

// mixed cutoffs for pairs not set explicitly (src/pair_GRM_LBDEM.cpp:636-646)
double PairLubricateGRMLBDEM::init_one(int i, int j)
{
  if (setflag[i][j] == 0) {
    cut_inner[i][j] = mix_distance(cut_inner[i][i], cut_inner[j][j]);
    cut[i][j] = mix_distance(cut[i][i], cut[j][j]);
  }
  cut_inner[j][i] = cut_inner[i][j];
  return cut[i][j];
}

// restart layout of src/pair_GRM_LBDEM.cpp:652-737: settings, then per type
// pair i <= j the flag and, when set, cut_inner and cut
void PairLubricateGRMLBDEM::write_restart(FILE *fp)
{
  write_restart_settings(fp);
  for (int i = 1; i <= atom->ntypes; i++)
    for (int j = i; j <= atom->ntypes; j++) {
      fwrite(&setflag[i][j], sizeof(int), 1, fp);
      if (setflag[i][j]) {
        fwrite(&cut_inner[i][j], sizeof(double), 1, fp);
        fwrite(&cut[i][j], sizeof(double), 1, fp);
      }
    }
}

void PairLubricateGRMLBDEM::read_restart(FILE *fp)
{
  read_restart_settings(fp);
  allocate();
  const bool root = comm->me == 0;
  for (int i = 1; i <= atom->ntypes; i++)
    for (int j = i; j <= atom->ntypes; j++) {
      if (root) utils::sfread(FLERR, &setflag[i][j], sizeof(int), 1, fp, nullptr, error);
      MPI_Bcast(&setflag[i][j], 1, MPI_INT, 0, world);
      if (!setflag[i][j]) continue;
      if (root) {
        utils::sfread(FLERR, &cut_inner[i][j], sizeof(double), 1, fp, nullptr, error);
        utils::sfread(FLERR, &cut[i][j], sizeof(double), 1, fp, nullptr, error);
      }
      MPI_Bcast(&cut_inner[i][j], 1, MPI_DOUBLE, 0, world);
      MPI_Bcast(&cut[i][j], 1, MPI_DOUBLE, 0, world);
    }
}

void PairLubricateGRMLBDEM::write_restart_settings(FILE *fp)
{
  fwrite(&mu, sizeof(double), 1, fp);
  fwrite(&cut_inner_global, sizeof(double), 1, fp);
  fwrite(&cut_global, sizeof(double), 1, fp);
  fwrite(&offset_flag, sizeof(int), 1, fp);
  fwrite(&mix_flag, sizeof(int), 1, fp);
  fwrite(&flagGRM, sizeof(int), 1, fp);
  fwrite(&cutoff_factor, sizeof(double), 1, fp);
}

void PairLubricateGRMLBDEM::read_restart_settings(FILE *fp)
{
  if (comm->me == 0) {
    utils::sfread(FLERR, &mu, sizeof(double), 1, fp, nullptr, error);
    utils::sfread(FLERR, &cut_inner_global, sizeof(double), 1, fp, nullptr, error);
    utils::sfread(FLERR, &cut_global, sizeof(double), 1, fp, nullptr, error);
    utils::sfread(FLERR, &offset_flag, sizeof(int), 1, fp, nullptr, error);
    utils::sfread(FLERR, &mix_flag, sizeof(int), 1, fp, nullptr, error);
    utils::sfread(FLERR, &flagGRM, sizeof(int), 1, fp, nullptr, error);
    utils::sfread(FLERR, &cutoff_factor, sizeof(double), 1, fp, nullptr, error);
  }
  MPI_Bcast(&mu, 1, MPI_DOUBLE, 0, world);
  MPI_Bcast(&cut_inner_global, 1, MPI_DOUBLE, 0, world);
  MPI_Bcast(&cut_global, 1, MPI_DOUBLE, 0, world);
  MPI_Bcast(&offset_flag, 1, MPI_INT, 0, world);
  MPI_Bcast(&mix_flag, 1, MPI_INT, 0, world);
  MPI_Bcast(&flagGRM, 1, MPI_INT, 0, world);
  MPI_Bcast(&cutoff_factor, 1, MPI_DOUBLE, 0, world);
}

// ghost v and omega, 6 values per atom (src/pair_GRM_LBDEM.cpp:741-778)
int PairLubricateGRMLBDEM::pack_comm(int n, int *list_, double *buf, int, int *)
{
  int m = 0;
  for (int i = 0; i < n; i++) {
    const int j = list_[i];
    for (int k = 0; k < 3; k++) buf[m++] = atom->v[j][k];
    for (int k = 0; k < 3; k++) buf[m++] = atom->omega[j][k];
  }
  return 6;
}

void PairLubricateGRMLBDEM::unpack_comm(int n, int first, double *buf)
{
  int m = 0;
  for (int i = first; i < first + n; i++) {
    for (int k = 0; k < 3; k++) atom->v[i][k] = buf[m++];
    for (int k = 0; k < 3; k++) atom->omega[i][k] = buf[m++];
  }
}

// fix adapt may change mu (src/pair_GRM_LBDEM.cpp:786-800)
int PairLubricateGRMLBDEM::pre_adapt(char *name, int, int, int, int)
{
  return strcmp(name, "mu") == 0 ? 0 : -1;
}

void PairLubricateGRMLBDEM::adapt(int, int, int, int, int, double value)
{
  mu = value;
}
