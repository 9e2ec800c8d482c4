// Host build of the library's solid-fraction routine, so the CPU tests can
// compare it with the oracle's point-by-point count.
#include "solid_fraction.h"

extern "C" double sf_fast(int dim, double hsq_half, int i, int j, int k, double xP, double yP,
                          double zP, double rP) {
  return lbmpsm::solid_fraction(dim, hsq_half, i, j, k, xP, yP, zP, rP);
}
