/* ----------------------------------------------------------------------
   pair_style lubricate/GRM/LBDEM: sub-grid lubrication correction for
   resolved LBM-DEM suspensions (replaces src/pair_GRM_LBDEM.{h,cpp}).

   For every pair of a full neighbour list whose surface gap lies below an
   outer gap h_N, one term of the Jeffrey-Onishi grand resistance matrix,
   chosen by flagGRM, is evaluated at the gap (clipped from below at the
   inner cutoff) minus the same term at h_N, and added to atom i only:
     1  squeeze force     X^A (normal relative velocity)
     2  shear force       Y^A (tangential velocities)
     3  torque from translation   Y^B (v x e)
     4  force from rotation       Y^B (e x omega)
     5  torque from rotation      Y^C (tangential angular velocities)

   Syntax (src/pair_GRM_LBDEM.cpp:502-546):
     pair_style lubricate/GRM/LBDEM mu cut_inner cut flagGRM cutoff_factor
     pair_coeff I J [cut_inner cut]
   h_N = cut_ij * cutoff_factor (a gap, not a centre distance).  Requires
   newton pair off, ghost velocities and atom_style sphere.
------------------------------------------------------------------------- */

#ifdef PAIR_CLASS
// clang-format off
PairStyle(lubricate/GRM/LBDEM,PairLubricateGRMLBDEM)
// clang-format on
#else

#ifndef LMP_PAIR_LUBRICATE_GRMLBDEM_H
#define LMP_PAIR_LUBRICATE_GRMLBDEM_H

#include <cstdio>

#include "pair.h"

namespace LAMMPS_NS {

class PairLubricateGRMLBDEM : public Pair {
 public:
  PairLubricateGRMLBDEM(class LAMMPS *);
  ~PairLubricateGRMLBDEM() override;
  void compute(int, int) override;
  void settings(int, char **) override;
  void coeff(int, char **) override;
  double init_one(int, int) override;
  void init_style() override;
  void write_restart(FILE *) override;
  void read_restart(FILE *) override;
  void write_restart_settings(FILE *);
  void read_restart_settings(FILE *);
  int pre_adapt(char *, int, int, int, int);
  void adapt(int, int, int, int, int, double);

  int pack_comm(int, int *, double *, int, int *);
  void unpack_comm(int, int, double *);

 protected:
  double mu = 0.0, cut_inner_global = 0.0, cut_global = 0.0;
  int flagGRM = 0;
  double cutoff_factor = 1.0;
  int shearing = 0, flagdeform = 0, flagwall = 0;

  double **cut_inner = nullptr, **cut = nullptr;

  void allocate();
};

}    // namespace LAMMPS_NS

#endif
#endif
