// Host build of the library's solid-fraction routines, so the CPU tests can
// compare them with the oracle's point-by-point count.
#include "solid_fraction.h"

extern "C" double sf_fast(int dim, double hsq_half, int i, int j, int k, double xP, double yP,
                          double zP, double rP) {
  return lbmpsm::solid_fraction(dim, hsq_half, i, j, k, xP, yP, zP, rP);
}

// The map kernel's split: classification, then `groups` lanes each counting
// the pairs g, g + groups, ... and the integer parts summed.
extern "C" double sf_split(int dim, double hsq_half, int i, int j, int k, double xP, double yP,
                           double zP, double rP, int groups) {
  const double dx = (double)i - xP, dy = (double)j - yP, dz = (double)k - zP;
  const int cls = lbmpsm::sf_classify(hsq_half, dx, dy, dz, rP);
  if (cls == lbmpsm::SF_ZERO) return 0.0;
  if (cls == lbmpsm::SF_ONE) return 1.0;
  int n = 0;
  for (int g = 0; g < groups; ++g)
    n += lbmpsm::sf_count_pairs(dim, dx, dy, dz, rP, g, lbmpsm::kSfPairs, groups);
  return lbmpsm::sf_from_count(dim, n);
}

// The map kernel's column form with a given column axis (-1: its choice).
extern "C" double sf_columns(int dim, double hsq_half, int i, int j, int k, double xP, double yP,
                             double zP, double rP, int axis) {
  const double dx = (double)i - xP, dy = (double)j - yP, dz = (double)k - zP;
  const int cls = lbmpsm::sf_classify(hsq_half, dx, dy, dz, rP);
  if (cls == lbmpsm::SF_ZERO) return 0.0;
  if (cls == lbmpsm::SF_ONE) return 1.0;
  return lbmpsm::sf_from_count(dim, lbmpsm::sf_count_columns(dim, dx, dy, dz, rP, axis));
}
