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

#include <math.h>

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

// ---------------------------------------------------------------------------
// Count by columns along a chosen axis (the map kernel's form).
//
// Every sub-point test is the reference's fl(fl(X + Y) + Z) < r^2 with
// X, Y, Z the squared sub-point offsets along x, y, z.  That value is
// monotone in each of X, Y, Z, so along any axis a column of 10 points is
// all-in when its largest square passes and all-out when its smallest fails;
// only the columns in between are tested point by point.  The column axis
// is free: choosing the one with the smallest |offset| (the axis most nearly
// tangent to the sphere at the cell) leaves the fewest undecided columns.
// The other two axes are the "lane" axis (one per lane in the kernel) and
// the "loop" axis:
//   column z: lane x, loop y     column x: lane y, loop z     column y: lane x, loop z
// ---------------------------------------------------------------------------

// Squared offsets of the 10 sub-point slices along one axis.
LBM_HD void sf_squares(double d, double* sq) {
  for (int c = 0; c < kSfSlices; ++c) {
    const double t = d + sf_delta(c);
    sq[c] = t * t;
  }
}

LBM_HD int sf_column_axis(double dx, double dy, double dz) {
  const double ax = dx < 0.0 ? -dx : dx, ay = dy < 0.0 ? -dy : dy, az = dz < 0.0 ? -dz : dz;
  if (az <= ax && az <= ay) return 2;
  if (ax <= ay) return 0;
  return 1;
}

// Offsets (dx, dy, dz) of the lane, loop and column axes for a column axis.
LBM_HD void sf_axis_offsets(int axis, double dx, double dy, double dz, double& lane,
                            double& loop, double& col) {
  if (axis == 2) {
    lane = dx; loop = dy; col = dz;
  } else if (axis == 0) {
    lane = dy; loop = dz; col = dx;
  } else {
    lane = dx; loop = dz; col = dy;
  }
}

// fl(fl(X + Y) + Z) from the lane, loop and column squares.
template <int AXIS>
LBM_HD double sf_point(double lane_sq, double loop_sq, double col_sq) {
  if constexpr (AXIS == 2) return (lane_sq + loop_sq) + col_sq;
  else if constexpr (AXIS == 0) return (col_sq + lane_sq) + loop_sq;
  else return (lane_sq + col_sq) + loop_sq;
}

// Decided columns of one lane slice: returns their count and marks the
// undecided loop indices in `undec`.
template <int AXIS>
LBM_HD int sf_lane_count(double lane_sq, const double* loop_sq, double col_min, double col_max,
                         double rSq, unsigned& undec) {
  int n = 0;
  undec = 0u;
  for (int m = 0; m < kSfSlices; ++m) {
    if (!(sf_point<AXIS>(lane_sq, loop_sq[m], col_min) < rSq)) continue;
    if (sf_point<AXIS>(lane_sq, loop_sq[m], col_max) < rSq) {
      n += kSfSlices;
      continue;
    }
    undec |= 1u << m;
  }
  return n;
}

template <int AXIS>
LBM_HD int sf_column_count(double lane_sq, double loop_sq, const double* col_sq, double rSq) {
  int n = 0;
  for (int c = 0; c < kSfSlices; ++c) n += sf_point<AXIS>(lane_sq, loop_sq, col_sq[c]) < rSq;
  return n;
}

// 2D: all 100 tests fl(X + Y) < r^2 of one lane slice.
LBM_HD int sf_lane_count_2d(double lane_sq, const double* loop_sq, double rSq) {
  int n = 0;
  for (int m = 0; m < kSfSlices; ++m) n += (lane_sq + loop_sq[m]) < rSq;
  return n;
}

// Whole-cell count in the column form, with the column axis given
// (axis < 0: the kernel's choice).  Host reference of the kernel's split.
template <int AXIS>
LBM_HD int sf_count_columns_axis(double lane, double loop, double col, double rSq) {
  double ls[kSfSlices], ms[kSfSlices], cs[kSfSlices];
  sf_squares(lane, ls);
  sf_squares(loop, ms);
  sf_squares(col, cs);
  double cmin = cs[0], cmax = cs[0];
  for (int c = 1; c < kSfSlices; ++c) {
    cmin = cs[c] < cmin ? cs[c] : cmin;
    cmax = cs[c] > cmax ? cs[c] : cmax;
  }
  int n = 0;
  for (int l = 0; l < kSfSlices; ++l) {
    unsigned undec = 0u;
    n += sf_lane_count<AXIS>(ls[l], ms, cmin, cmax, rSq, undec);
    for (int m = 0; m < kSfSlices; ++m)
      if (undec & (1u << m)) n += sf_column_count<AXIS>(ls[l], ms[m], cs, rSq);
  }
  return n;
}

LBM_HD int sf_count_columns(int dim, double dx, double dy, double dz, double rP_LB, int axis) {
  const double rSq = rP_LB * rP_LB;
  if (dim == 2) {
    double xs[kSfSlices], ys[kSfSlices];
    sf_squares(dx, xs);
    sf_squares(dy, ys);
    int n = 0;
    for (int l = 0; l < kSfSlices; ++l) n += sf_lane_count_2d(xs[l], ys, rSq);
    return n;
  }
  if (axis < 0) axis = sf_column_axis(dx, dy, dz);
  double lane, loop, col;
  sf_axis_offsets(axis, dx, dy, dz, lane, loop, col);
  if (axis == 2) return sf_count_columns_axis<2>(lane, loop, col, rSq);
  if (axis == 0) return sf_count_columns_axis<0>(lane, loop, col, rSq);
  return sf_count_columns_axis<1>(lane, loop, col, rSq);
}

// ---------------------------------------------------------------------------
// Count by lines (the 3D map kernel's form).
//
// Fix the x and y sub-point offsets (a, b) of a column of cells (i, j) and
// run along z over every cell k of a zone and every slice c: the reference
// tests fl(s + Z_p) < r^2 with s = fl(X_a + Y_b) and Z_p the squared z
// offset of point p = 10 k + c.  The offsets fl(fl(k - z) + delta_c) grow
// with p, so Z_p falls and then rises (a "V"), and the passing points form
// one interval around the minimum.  Two binary searches find it; every
// cell of the column then receives the length of its overlap.
// ---------------------------------------------------------------------------

// Square root for the end-point prediction only (the exact tests decide):
// the hardware approximation on the GPU.
LBM_HD float sf_sqrt_approx(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_sqrtf(x);
#elif defined(LBMPSM_SF_HOST_SQRT)
  return LBMPSM_SF_HOST_SQRT(x);  // a host test's square root with a set error
#else
  return sqrtf(x);
#endif
}

// First index in [lo, hi] that passes, given that passing is monotone
// non-decreasing on [lo, hi] and hi passes.
LBM_HD int sf_first_pass(const double* Z, int lo, int hi, double s, double rSq) {
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (s + Z[mid] < rSq) hi = mid;
    else lo = mid + 1;
  }
  return lo;
}

// Last index in [lo, hi] that passes, given that passing is monotone
// non-increasing on [lo, hi] and lo passes.
LBM_HD int sf_last_pass(const double* Z, int lo, int hi, double s, double rSq) {
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (s + Z[mid] < rSq) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// The interval of passing points of one line.  The points' z offsets are
// t_p = t_0 + p/10 (up to ~1e-14), and a point passes the exact test
// s + t_p^2 < r^2 when |t_p| < h = sqrt(r^2 - s) (the test's rounding moves
// that bound by ~1e-14 of a point).  Both ends are predicted in float:
// v_lo = 10 (-h - t_0) and v_hi = 10 (h - t_0); the first passing point is
// floor(v_lo) + 1, the last floor(v_hi).  With e = 2^-24, the float values
// differ from the real ones by at most 10 (h + |t_0|) e (2k + 3.5) for a
// square root within k ulp (k = 2: 75 e (h + |t_0|)).  When v lies further
// than tol = (h + |t_0|) 2^-16 (> 3x that bound) from an integer, its floor
// is the exact one and the end needs no FP64 test.  Otherwise the end is
// confirmed by its two exact tests (the end passes, its outer neighbour
// fails), and failing those the binary search decides.  Returns false for
// an empty line.
LBM_HD bool sf_line_interval(const double* Z, int P, int pstar, double t0, double s, double rSq,
                             int& lo, int& hi) {
  if (!(s + Z[pstar] < rSq)) return false;
  const float h = sf_sqrt_approx((float)(rSq - s));
  const float x0 = (float)t0;
  const float tol = (h + fabsf(x0)) * (1.0f / 65536.0f);
  const float vl = (-h - x0) * 10.0f, fl = floorf(vl);
  int pe = (int)fl + 1;
  pe = pe < 0 ? 0 : (pe > pstar ? pstar : pe);
  const float rl = vl - fl;
  if ((rl > tol && rl < 1.0f - tol) || (s + Z[pe] < rSq && (pe == 0 || !(s + Z[pe - 1] < rSq))))
    lo = pe;
  else
    lo = sf_first_pass(Z, 0, pstar, s, rSq);
  const float vh = (h - x0) * 10.0f, fh = floorf(vh);
  pe = (int)fh;
  pe = pe < pstar ? pstar : (pe > P - 1 ? P - 1 : pe);
  const float rh = vh - fh;
  if ((rh > tol && rh < 1.0f - tol) || (s + Z[pe] < rSq && (pe == P - 1 || !(s + Z[pe + 1] < rSq))))
    hi = pe;
  else
    hi = sf_last_pass(Z, pstar, P - 1, s, rSq);
  return true;
}

// Index of the smallest square along a line (first one among equals).
LBM_HD int sf_line_min(const double* Z, int P) {
  int m = 0;
  for (int p = 1; p < P; ++p)
    if (Z[p] < Z[m]) m = p;
  return m;
}

// Cells [klo, khi] of a line receive its passing points [lo, hi]: the two
// end cells directly (part), the cells in between 10 each (diff, summed
// along the line afterwards).
LBM_HD void sf_line_split(int lo, int hi, int& klo, int& khi, int& nlo, int& nhi) {
  klo = lo / kSfSlices;
  khi = hi / kSfSlices;
  if (klo == khi) {
    nlo = hi - lo + 1;
    nhi = 0;
  } else {
    nlo = kSfSlices * klo + kSfSlices - lo;
    nhi = hi - kSfSlices * khi + 1;
  }
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
