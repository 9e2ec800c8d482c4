// Host build of the library's solid-fraction routines, so the CPU tests can
// compare them with the oracle's point-by-point count.  Built a second time
// with LBMPSM_SF_HOST_SQRT: the end-point prediction's float square root is
// then off by a set number of ulps, larger than the GPU's, and the line form
// must still give the exact counts.
#ifdef LBMPSM_SF_HOST_SQRT
#include <math.h>
#include <stdlib.h>
static int g_sqrt_ulps = 0;
static float sf_sqrt_off(float x) {
  float r = sqrtf(x);
  for (int u = 0; u < abs(g_sqrt_ulps); ++u) r = nextafterf(r, g_sqrt_ulps > 0 ? INFINITY : 0.0f);
  return r;
}
extern "C" void sf_set_sqrt_ulps(int k) { g_sqrt_ulps = k; }
#endif
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

#include <vector>

// Counts of every cell of an nj x nk plane of a zone by the line form,
// checked against the per-cell count (sf_count_pairs) for every cell that
// needs counting.  dx is the plane's x offset, dy0/dz0 the offsets of the
// first column / cell are given through the reference formula
// (double)(J) - yP with J = j0 + j.  Returns the number of mismatches.
extern "C" int sf_lines_plane(double hsq_half, int i, int j0, int k0, int nj, int nk, double xP,
                              double yP, double zP, double rP) {
  using namespace lbmpsm;
  const double rSq = rP * rP;
  const double dx = (double)i - xP;
  std::vector<double> X(kSfSlices), Y(kSfSlices * nj), Z(kSfSlices * nk);
  sf_squares(dx, X.data());
  for (int j = 0; j < nj; ++j) sf_squares((double)(j0 + j) - yP, &Y[kSfSlices * j]);
  for (int k = 0; k < nk; ++k) sf_squares((double)(k0 + k) - zP, &Z[kSfSlices * k]);
  const int P = kSfSlices * nk;
  const int pstar = sf_line_min(Z.data(), P);
  std::vector<int> part(nj * (nk + 1), 0), diff(nj * (nk + 1), 0);
  for (int j = 0; j < nj; ++j)
    for (int a = 0; a < kSfSlices; ++a)
      for (int b = 0; b < kSfSlices; ++b) {
        const double s = X[a] + Y[kSfSlices * j + b];
        int lo, hi;
        if (!sf_line_interval(Z.data(), P, pstar, (double)k0 - zP + sf_delta(0), s, rSq, lo, hi))
          continue;
        if (s + Z[pstar] < rSq) {
          // the predicted ends agree with the plain binary searches
          if (lo != sf_first_pass(Z.data(), 0, pstar, s, rSq)) return -1;
          if (hi != sf_last_pass(Z.data(), pstar, P - 1, s, rSq)) return -1;
        }
        int klo, khi, nlo, nhi;
        sf_line_split(lo, hi, klo, khi, nlo, nhi);
        part[j * (nk + 1) + klo] += nlo;
        if (khi > klo) {
          part[j * (nk + 1) + khi] += nhi;
          if (khi > klo + 1) {
            diff[j * (nk + 1) + klo + 1] += kSfSlices;
            diff[j * (nk + 1) + khi] -= kSfSlices;
          }
        }
      }
  int bad = 0;
  for (int j = 0; j < nj; ++j) {
    int run = 0;
    for (int k = 0; k < nk; ++k) {
      run += diff[j * (nk + 1) + k];
      const double dy = (double)(j0 + j) - yP, dz = (double)(k0 + k) - zP;
      if (sf_classify(hsq_half, dx, dy, dz, rP) != SF_COUNT) continue;
      const int want = sf_count_pairs(3, dx, dy, dz, rP, 0, kSfPairs, 1);
      if (part[j * (nk + 1) + k] + run != want) ++bad;
    }
  }
  return bad;
}
