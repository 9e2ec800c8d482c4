// Sub-sampled solid fraction of a lattice cell covered by a sphere (3D) or
// disk (2D), src/LBM_PSM_exchangeParticleData.cpp:116-169.
//
// The reference classifies the cell by its centre distance (clearly outside
// -> 0, clearly inside -> 1) and otherwise counts the sub-points
//   (a, b, c), a,b,c in 0..9:  fl(fl(dx_sq[a] + dy_sq[b]) + dz_sq[c]) < r^2
// (2D: fl(dx_sq[a] + dy_sq[b]) < r^2) and returns fraction * count.
//
// Here the count is split over the 100 (a, b) pairs so that several lanes
// can share one cell; every pair's count is the reference's own set of tests,
// so any split sums to the reference's integer.  Within a pair, rounded
// addition is monotone in each operand, so the smallest and largest dz_sq
// decide the whole c column when both agree; only undecided columns are
// tested point by point.
//
// Compiles as HIP device code and as plain host C++ (the CPU tests check it
// against the oracle's point-by-point count).
#pragma once

#if defined(__HIPCC__)
#define LBM_HD __host__ __device__ __forceinline__
#else
#define LBM_HD inline
#endif

namespace lbmpsm {

constexpr int kSfSlices = 10;
constexpr int kSfPairs = kSfSlices * kSfSlices;

enum SfClass : int { SF_ZERO = 0, SF_ONE = 1, SF_COUNT = 2 };

// Early exits of calcSolidFraction (:130-140).  hsq_half is sqrt(2.1)/2
// (2D) or sqrt(3.1)/2 (3D) as the reference computes it.
LBM_HD int sf_classify(double hsq_half, double dx, double dy, double dz, double rP_LB) {
  double dist = dx * dx + dy * dy + dz * dz;
  double rP = rP_LB + hsq_half;
  if (dist > rP * rP) return SF_ZERO;
  double rM = rP_LB - hsq_half;
  if (dist < rM * rM) return SF_ONE;
  return SF_COUNT;
}

// Offset of slice a from the node: -0.5 + (a + 0.5) * (1/10), :145.
LBM_HD double sf_delta(int a) {
  const double sliceWidth = 1.0 / ((double)kSfSlices);
  return -0.5 + ((double)a + 0.5) * sliceWidth;
}

// Count of the reference's sub-point tests that pass, over the (a,b) pairs
// p = p0, p0 + stride, ... < p1 (p = 10 a + b).
LBM_HD int sf_count_pairs(int dim, double dx, double dy, double dz, double rP_LB, int p0, int p1,
                          int stride) {
  const double rSq = rP_LB * rP_LB;
  int n = 0;
  if (dim == 2) {
    for (int p = p0; p < p1; p += stride) {
      const double dxi = dx + sf_delta(p / kSfSlices);
      const double dyi = dy + sf_delta(p % kSfSlices);
      if (dxi * dxi + dyi * dyi < rSq) ++n;
    }
    return n;
  }
  double zs[kSfSlices];
  double zmin = 1e308, zmax = -1.0;
  for (int c = 0; c < kSfSlices; ++c) {
    const double dzi = dz + sf_delta(c);
    zs[c] = dzi * dzi;
    zmin = zs[c] < zmin ? zs[c] : zmin;
    zmax = zs[c] > zmax ? zs[c] : zmax;
  }
  for (int p = p0; p < p1; p += stride) {
    const double dxi = dx + sf_delta(p / kSfSlices);
    const double dyi = dy + sf_delta(p % kSfSlices);
    const double s = dxi * dxi + dyi * dyi;
    if (!(s + zmin < rSq)) continue;
    if (s + zmax < rSq) {
      n += kSfSlices;
      continue;
    }
    for (int c = 0; c < kSfSlices; ++c)
      if (s + zs[c] < rSq) ++n;
  }
  return n;
}

// fraction * count, :120-124 and :168.
LBM_HD double sf_from_count(int dim, int n) {
  double fraction = 1.0 / ((double)(kSfSlices * kSfSlices));
  if (dim == 3) fraction = 1.0 / ((double)(kSfSlices * kSfSlices * kSfSlices));
  return fraction * ((double)n);
}

// The whole calcSolidFraction for node (i,j,k) (reference halo'd indices)
// and a particle at (xP, yP, zP) with radius rP_LB.
LBM_HD double solid_fraction(int dim, double hsq_half, int i, int j, int k, double xP,
                             double yP, double zP, double rP_LB) {
  const double dx = (double)i - xP;
  const double dy = (double)j - yP;
  const double dz = (double)k - zP;
  const int cls = sf_classify(hsq_half, dx, dy, dz, rP_LB);
  if (cls == SF_ZERO) return 0.0;
  if (cls == SF_ONE) return 1.0;
  return sf_from_count(dim, sf_count_pairs(dim, dx, dy, dz, rP_LB, 0, kSfPairs, 1));
}

}  // namespace lbmpsm
