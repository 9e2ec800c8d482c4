// Sub-sampled solid fraction of a lattice cell covered by a sphere (3D) or
// disk (2D), src/LBM_PSM_exchangeParticleData.cpp:116-169.
//
// The reference tests all 10^dim sub-points; the count here is the same
// number obtained with fewer tests.  Each sub-point test is
//   fl(fl(dx_sq[a] + dy_sq[b]) + dz_sq[c]) < r^2
// and rounded addition is monotone in each operand, so whole rows (fixed a)
// or columns (fixed a,b) can be decided from the smallest and largest of the
// remaining squares: if even the smallest fails, all fail; if even the
// largest passes, all pass.  Only the undecided columns are tested point by
// point.  The result is the reference's count, not an approximation of it.
//
// Compiles as HIP device code and as plain host C++ (for the CPU test that
// checks it against the point-by-point count).
#pragma once

#if defined(__HIPCC__)
#define LBM_HD __host__ __device__ __forceinline__
#else
#define LBM_HD inline
#endif

namespace lbmpsm {

// sqrt(2.1)/2 and sqrt(3.1)/2 as the reference computes them; passed in by
// the caller so host and device use the same doubles.
LBM_HD double solid_fraction(int dim, double hsq_half, int i, int j, int k, double xP,
                             double yP, double zP, double rP_LB) {
  const int slices = 10;
  const double sliceWidth = 1.0 / ((double)slices);
  double fraction = (dim == 3) ? 1.0 / ((double)(slices * slices * slices))
                               : 1.0 / ((double)(slices * slices));
  double dx = (double)i - xP;
  double dy = (double)j - yP;
  double dz = (double)k - zP;
  double dist = dx * dx + dy * dy + dz * dz;
  double rP = rP_LB + hsq_half;
  if (dist > rP * rP) return 0.0;
  double rM = rP_LB - hsq_half;
  if (dist < rM * rM) return 1.0;
  double rSq = rP_LB * rP_LB;

  double xs[10], ys[10], zs[10];
  double ymin = 1e308, ymax = -1.0, zmin = 1e308, zmax = -1.0;
  for (int a = 0; a < slices; ++a) {
    double delta = -0.5 + ((double)a + 0.5) * sliceWidth;
    double dxi = dx + delta;
    double dyi = dy + delta;
    double dzi = dz + delta;
    xs[a] = dxi * dxi;
    ys[a] = dyi * dyi;
    zs[a] = dzi * dzi;
    ymin = ys[a] < ymin ? ys[a] : ymin;
    ymax = ys[a] > ymax ? ys[a] : ymax;
    zmin = zs[a] < zmin ? zs[a] : zmin;
    zmax = zs[a] > zmax ? zs[a] : zmax;
  }
  int n = 0;
  if (dim == 2) {
    for (int a = 0; a < slices; ++a) {
      if (!(xs[a] + ymin < rSq)) continue;
      if (xs[a] + ymax < rSq) { n += slices; continue; }
      for (int b = 0; b < slices; ++b)
        if (xs[a] + ys[b] < rSq) ++n;
    }
  } else {
    for (int a = 0; a < slices; ++a) {
      if (!((xs[a] + ymin) + zmin < rSq)) continue;
      if ((xs[a] + ymax) + zmax < rSq) { n += slices * slices; continue; }
      for (int b = 0; b < slices; ++b) {
        double s = xs[a] + ys[b];
        if (!(s + zmin < rSq)) continue;
        if (s + zmax < rSq) { n += slices; continue; }
        for (int c = 0; c < slices; ++c)
          if (s + zs[c] < rSq) ++n;
      }
    }
  }
  return fraction * ((double)n);
}

}  // namespace lbmpsm
