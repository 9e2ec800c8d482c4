// Device-side building blocks shared by the lattice kernels: velocity sets,
// equilibrium, Guo forcing, macroscopic moments and the Zou-He faces, all
// written for a node whose q populations sit in registers.
//
// The FP64 forms keep the reference's operand order (file:line given per
// function) and the translation units that use them are compiled with
// -ffp-contract=off, so each value is the double the reference computes.
// Terms whose lattice-vector factor is 0 are left out: they only ever add a
// signed zero, which does not change a non-zero sum.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lbmpsm {

// Velocity sets, src/LBM_PSM_lattice.h:186-209 (order is load-bearing:
// the Zou-He faces address populations by these indices).
template <int Q>
struct VelSet;

template <>
struct VelSet<9> {
  static constexpr int ex[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
  static constexpr int ey[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
  static constexpr int ez[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  static constexpr double w[9] = {4.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0, 1.0 / 9.0,
                                  1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};
};

template <>
struct VelSet<19> {
  static constexpr int ex[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0};
  static constexpr int ey[19] = {0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1};
  static constexpr int ez[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, -1, 1, -1, 1};
  static constexpr double w[19] = {1.0 / 3.0,  1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0,
                                   1.0 / 18.0, 1.0 / 18.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
                                   1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
                                   1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};
};

// e_q . v with the zero components of e_q dropped, in x, y, z order.
template <int c, typename T>
__device__ __forceinline__ T eterm(T v) {
  if constexpr (c > 0) return v;
  else return -v;
}
template <int Q, int q, typename T>
__device__ __forceinline__ T edot(T vx, T vy, T vz) {
  constexpr int cx = VelSet<Q>::ex[q], cy = VelSet<Q>::ey[q], cz = VelSet<Q>::ez[q];
  if constexpr (cx != 0) {
    T s = eterm<cx>(vx);
    if constexpr (cy != 0) s = s + eterm<cy>(vy);
    if constexpr (cz != 0) s = s + eterm<cz>(vz);
    return s;
  } else if constexpr (cy != 0) {
    T s = eterm<cy>(vy);
    if constexpr (cz != 0) s = s + eterm<cz>(vz);
    return s;
  } else if constexpr (cz != 0) {
    return eterm<cz>(vz);
  } else {
    return T(0);
  }
}

// Compile-time loop helper: f(std::integral_constant-like index).
template <int I, int N>
struct Unroll {
  template <typename F>
  __device__ __forceinline__ static void run(F&& f) {
    f.template operator()<I>();
    Unroll<I + 1, N>::run(f);
  }
};
template <int N>
struct Unroll<N, N> {
  template <typename F>
  __device__ __forceinline__ static void run(F&&) {}
};

// Macroscopic moments, src/LBM_PSM_BGK_dynamics.cpp:48-77:
// rho = sum f (q order), j = sum f e, u = (j + F/2) / rho.
template <int Q, typename T>
__device__ __forceinline__ void moments(const T* f, T& rho, T& jx, T& jy, T& jz) {
  T r = T(0) + f[0];
  T x = T(0), y = T(0), z = T(0);
  Unroll<1, Q>::run([&]<int q>() {
    r = r + f[q];
    constexpr int cx = VelSet<Q>::ex[q], cy = VelSet<Q>::ey[q], cz = VelSet<Q>::ez[q];
    if constexpr (cx > 0) x = x + f[q];
    if constexpr (cx < 0) x = x - f[q];
    if constexpr (cy > 0) y = y + f[q];
    if constexpr (cy < 0) y = y - f[q];
    if constexpr (cz > 0) z = z + f[q];
    if constexpr (cz < 0) z = z - f[q];
  });
  rho = r;
  jx = x;
  jy = y;
  jz = z;
}

// ---------------------------------------------------------------------------
// Zou-He faces, src/LBM_PSM_zou_he_BC.cpp.  `r0` is added in front of every
// density sum; it is 0 for raw populations and rho0 for populations stored
// as f - w*rho0 (for which every other term of these formulas is unchanged
// because opposite directions carry equal weights).
// ---------------------------------------------------------------------------
enum FaceKind : int {
  FK_2D_VEL_XN = 0, FK_2D_VEL_XP, FK_2D_DENS_XP, FK_2D_VEL_YN, FK_2D_VEL_YP,
  FK_3D_VEL_XN, FK_3D_VEL_XP, FK_3D_DENS_XP, FK_3D_VEL_YN, FK_3D_VEL_YP,
  FK_3D_VEL_ZN, FK_3D_VEL_ZP
};

struct Face {
  int kind;
  int axis;      // axis normal to the face (0 x, 1 y, 2 z)
  int plane;     // stored coordinate of the face along `axis`
  int a_lo, a_hi;  // inclusive stored range along the first tangential axis
  int b_lo, b_hi;  // inclusive stored range along the second tangential axis
  double u[3];     // imposed velocity (velocity faces)
  double rho;      // imposed density (density faces)
};

template <typename T>
struct ZHConst {
  static constexpr T third = T(1.0) / T(3.0);
  static constexpr T twoThirds = T(2.0) / T(3.0);
  static constexpr T sixth = T(1.0) / T(6.0);
};

// Outlet velocity of a density face on shifted populations g = f - w*rho0
// (FP32 arithmetic; r0 = rho0 is the weight sum of the populations taken).
// The reference's -1 + (r0 + sum g)/rho_bc subtracts two numbers near 1 and
// keeps only ~6e-8 of an outlet velocity of order 1e-3 (about 1e-5 relative,
// with the same sign every step at a steady outlet); (r0 - rho_bc) is exact
// for close densities, so ((r0 - rho_bc) + sum g)/rho_bc is the same
// quantity without the cancellation.  The FP64 forms keep the reference's.
__device__ __forceinline__ float outlet_velocity_shifted(float rho_bc, float r0, float g_sum) {
  return ((r0 - rho_bc) + g_sum) / rho_bc;
}

// :25-40
template <typename T>
__device__ __forceinline__ void zh2_vel_xn(T* f, T ux, T r0) {
  const T s6 = ZHConst<T>::sixth, t23 = ZHConst<T>::twoThirds;
  T rho = T(1.0) / (T(1.0) - ux) * (r0 + f[0] + f[2] + f[4] + T(2.0) * (f[3] + f[6] + f[7]));
  f[5] = f[7] + T(0.5) * (f[4] - f[2]) + rho * ux * s6;
  f[8] = f[6] + T(0.5) * (f[2] - f[4]) + rho * ux * s6;
  f[1] = f[3] + t23 * rho * ux;
}
// :43-58
template <typename T>
__device__ __forceinline__ void zh2_vel_xp(T* f, T ux, T r0) {
  const T s6 = ZHConst<T>::sixth, t23 = ZHConst<T>::twoThirds;
  T rho = T(1.0) / (T(1.0) - ux) * (r0 + f[0] + f[2] + f[4] + T(2.0) * (f[1] + f[5] + f[8]));
  f[7] = f[5] + T(0.5) * (f[2] - f[4]) - rho * ux * s6;
  f[6] = f[8] + T(0.5) * (f[4] - f[2]) - rho * ux * s6;
  f[3] = f[1] - t23 * rho * ux;
}
// :61-76
template <typename T>
__device__ __forceinline__ void zh2_dens_xp(T* f, T rho_bc, T r0) {
  const T s6 = ZHConst<T>::sixth, t23 = ZHConst<T>::twoThirds;
  T ux;
  if constexpr (sizeof(T) == 8)
    ux = T(-1.0) + T(1.0) / rho_bc * (r0 + f[0] + f[2] + f[4] + T(2.0) * (f[1] + f[5] + f[8]));
  else
    ux = outlet_velocity_shifted(rho_bc, r0, f[0] + f[2] + f[4] + T(2.0) * (f[1] + f[5] + f[8]));
  f[6] = f[8] + T(0.5) * (f[4] - f[2]) - rho_bc * ux * s6;
  f[7] = f[5] + T(0.5) * (f[2] - f[4]) - rho_bc * ux * s6;
  f[3] = f[1] - t23 * rho_bc * ux;
}
// :79-94
template <typename T>
__device__ __forceinline__ void zh2_vel_yn(T* f, T ux, T r0) {
  T rho = r0 + f[0] + f[3] + f[1] + T(2.0) * (f[8] + f[4] + f[7]);
  f[6] = f[8] + T(0.5) * (f[1] - f[3]) - rho * ux * T(0.5);
  f[5] = f[7] + T(0.5) * (f[3] - f[1]) + rho * ux * T(0.5);
  f[2] = f[4];
}
// :97-112
template <typename T>
__device__ __forceinline__ void zh2_vel_yp(T* f, T ux, T r0) {
  T rho = r0 + f[0] + f[3] + f[1] + T(2.0) * (f[6] + f[2] + f[5]);
  f[7] = f[5] + T(0.5) * (f[1] - f[3]) - rho * ux * T(0.5);
  f[8] = f[6] + T(0.5) * (f[3] - f[1]) + rho * ux * T(0.5);
  f[4] = f[2];
}
// :153-185
template <typename T>
__device__ __forceinline__ void zh3_vel_xn(T* f, T ux, T uy, T uz, T r0) {
  const T t3 = ZHConst<T>::third, s6 = ZHConst<T>::sixth;
  T rho = T(1.0) / (T(1.0) - ux) *
          (r0 + f[0] + f[3] + f[4] + f[5] + f[6] + f[11] + f[17] + f[18] + f[12] +
           T(2.0) * (f[2] + f[14] + f[8] + f[16] + f[10]));
  T Nyx = T(0.5) * (f[3] + f[11] + f[17] - (f[4] + f[18] + f[12])) - t3 * rho * uy;
  T Nzx = T(0.5) * (f[5] + f[18] + f[11] - (f[6] + f[17] + f[12])) - t3 * rho * uz;
  f[1] = f[2] + t3 * rho * ux;
  f[13] = f[14] + rho * s6 * (ux - uy) + Nyx;
  f[7] = f[8] + rho * s6 * (ux + uy) - Nyx;
  f[9] = f[10] + rho * s6 * (ux + uz) - Nzx;
  f[15] = f[16] + rho * s6 * (ux - uz) + Nzx;
}
// :189-221
template <typename T>
__device__ __forceinline__ void zh3_vel_xp(T* f, T ux, T uy, T uz, T r0) {
  const T t3 = ZHConst<T>::third, s6 = ZHConst<T>::sixth;
  T rho = T(1.0) / (T(1.0) + ux) *
          (r0 + f[0] + f[3] + f[4] + f[5] + f[6] + f[11] + f[17] + f[18] + f[12] +
           T(2.0) * (f[1] + f[7] + f[13] + f[9] + f[15]));
  T Nyx = T(0.5) * (f[3] + f[11] + f[17] - (f[4] + f[18] + f[12])) - t3 * rho * uy;
  T Nzx = T(0.5) * (f[5] + f[18] + f[11] - (f[6] + f[17] + f[12])) - t3 * rho * uz;
  f[2] = f[1] - t3 * rho * ux;
  f[14] = f[13] + rho * s6 * (-ux + uy) - Nyx;
  f[8] = f[7] + rho * s6 * (-ux - uy) + Nyx;
  f[10] = f[9] + rho * s6 * (-ux - uz) + Nzx;
  f[16] = f[15] + rho * s6 * (-ux + uz) - Nzx;
}
// :225-257
template <typename T>
__device__ __forceinline__ void zh3_dens_xp(T* f, T rho_bc, T uy, T uz, T r0) {
  const T t3 = ZHConst<T>::third, s6 = ZHConst<T>::sixth;
  T ux;
  if constexpr (sizeof(T) == 8)
    ux = T(-1.0) + T(1.0) / rho_bc *
                       (r0 + f[0] + f[3] + f[4] + f[5] + f[6] + f[11] + f[17] + f[18] + f[12] +
                        T(2.0) * (f[1] + f[7] + f[13] + f[9] + f[15]));
  else
    ux = outlet_velocity_shifted(rho_bc, r0,
                                 f[0] + f[3] + f[4] + f[5] + f[6] + f[11] + f[17] + f[18] +
                                     f[12] + T(2.0) * (f[1] + f[7] + f[13] + f[9] + f[15]));
  T Nyx = T(0.5) * (f[3] + f[11] + f[17] - (f[4] + f[18] + f[12])) - t3 * rho_bc * uy;
  T Nzx = T(0.5) * (f[5] + f[18] + f[11] - (f[6] + f[17] + f[12])) - t3 * rho_bc * uz;
  f[2] = f[1] - t3 * rho_bc * ux;
  f[14] = f[13] + rho_bc * s6 * (-ux + uy) - Nyx;
  f[8] = f[7] + rho_bc * s6 * (-ux - uy) + Nyx;
  f[10] = f[9] + rho_bc * s6 * (-ux - uz) + Nzx;
  f[16] = f[15] + rho_bc * s6 * (-ux + uz) - Nzx;
}
// :261-293
template <typename T>
__device__ __forceinline__ void zh3_vel_yn(T* f, T ux, T uy, T uz, T r0) {
  const T t3 = ZHConst<T>::third, s6 = ZHConst<T>::sixth;
  T rho = T(1.0) / (T(1.0) - uy) *
          (r0 + f[0] + f[1] + f[2] + f[5] + f[6] + f[9] + f[15] + f[16] + f[10] +
           T(2.0) * (f[4] + f[13] + f[8] + f[18] + f[12]));
  T Nxy = T(0.5) * (f[1] + f[9] + f[15] - (f[2] + f[16] + f[10])) - t3 * rho * ux;
  T Nzy = T(0.5) * (f[5] + f[9] + f[16] - (f[6] + f[15] + f[10])) - t3 * rho * uz;
  f[3] = f[4] + t3 * rho * uy;
  f[7] = f[8] + rho * s6 * (uy + ux) - Nxy;
  f[14] = f[13] + rho * s6 * (uy - ux) + Nxy;
  f[11] = f[12] + rho * s6 * (uy + uz) - Nzy;
  f[17] = f[18] + rho * s6 * (uy - uz) + Nzy;
}
// :296-328
template <typename T>
__device__ __forceinline__ void zh3_vel_yp(T* f, T ux, T uy, T uz, T r0) {
  const T t3 = ZHConst<T>::third, s6 = ZHConst<T>::sixth;
  T rho = T(1.0) / (T(1.0) + uy) *
          (r0 + f[0] + f[1] + f[2] + f[5] + f[6] + f[9] + f[15] + f[16] + f[10] +
           T(2.0) * (f[3] + f[7] + f[14] + f[11] + f[17]));
  T Nxy = T(0.5) * (f[1] + f[9] + f[15] - (f[2] + f[16] + f[10])) - t3 * rho * ux;
  T Nzy = T(0.5) * (f[5] + f[9] + f[16] - (f[6] + f[15] + f[10])) - t3 * rho * uz;
  f[4] = f[3] - t3 * rho * uy;
  f[8] = f[7] + rho * s6 * (-uy - ux) + Nxy;
  f[13] = f[14] + rho * s6 * (-uy + ux) - Nxy;
  f[12] = f[11] + rho * s6 * (-uy - uz) + Nzy;
  f[18] = f[17] + rho * s6 * (-uy + uz) - Nzy;
}
// :331-363
template <typename T>
__device__ __forceinline__ void zh3_vel_zn(T* f, T ux, T uy, T uz, T r0) {
  const T t3 = ZHConst<T>::third, s6 = ZHConst<T>::sixth;
  T rho = T(1.0) / (T(1.0) - uz) *
          (r0 + f[0] + f[1] + f[2] + f[3] + f[4] + f[7] + f[13] + f[14] + f[8] +
           T(2.0) * (f[6] + f[15] + f[10] + f[17] + f[12]));
  T Nxz = T(0.5) * (f[1] + f[7] + f[13] - (f[2] + f[14] + f[8])) - t3 * rho * ux;
  T Nyz = T(0.5) * (f[3] + f[7] + f[14] - (f[4] + f[13] + f[8])) - t3 * rho * uy;
  f[5] = f[6] + t3 * rho * uz;
  f[9] = f[10] + rho * s6 * (uz + ux) - Nxz;
  f[16] = f[15] + rho * s6 * (uz - ux) + Nxz;
  f[11] = f[12] + rho * s6 * (uz + uy) - Nyz;
  f[18] = f[17] + rho * s6 * (uz - uy) + Nyz;
}
// :366-398
template <typename T>
__device__ __forceinline__ void zh3_vel_zp(T* f, T ux, T uy, T uz, T r0) {
  const T t3 = ZHConst<T>::third, s6 = ZHConst<T>::sixth;
  T rho = T(1.0) / (T(1.0) + uz) *
          (r0 + f[0] + f[1] + f[2] + f[3] + f[4] + f[7] + f[14] + f[8] + f[13] +
           T(2.0) * (f[5] + f[9] + f[16] + f[11] + f[18]));
  T Nxz = T(0.5) * (f[1] + f[7] + f[13] - (f[2] + f[14] + f[8])) - t3 * rho * ux;
  T Nyz = T(0.5) * (f[3] + f[7] + f[14] - (f[4] + f[13] + f[8])) - t3 * rho * uy;
  f[6] = f[5] - t3 * rho * uz;
  f[15] = f[16] + rho * s6 * (-uz + ux) - Nxz;
  f[10] = f[9] + rho * s6 * (-uz - ux) + Nxz;
  f[17] = f[18] + rho * s6 * (-uz + uy) - Nyz;
  f[12] = f[11] + rho * s6 * (-uz - uy) + Nyz;
}

// Apply the faces whose plane and tangential ranges contain node (i,j,k),
// in table order; later faces see the values written by earlier ones, as
// the reference's face-by-face sweep does (src/fix_LBM_PSM_BC.cpp:223-401).
template <int Q, typename T>
__device__ __forceinline__ void apply_faces(T* f, const Face* faces, int nfaces, int i, int j,
                                            int k, T r0) {
  for (int fi = 0; fi < nfaces; ++fi) {
    const Face& F = faces[fi];
    int c = F.axis == 0 ? i : (F.axis == 1 ? j : k);
    if (c != F.plane) continue;
    int a = F.axis == 0 ? j : i;
    int b = F.axis == 2 ? j : k;
    if (a < F.a_lo || a > F.a_hi || b < F.b_lo || b > F.b_hi) continue;
    const T ux = T(F.u[0]), uy = T(F.u[1]), uz = T(F.u[2]);
    if constexpr (Q == 9) {
      switch (F.kind) {
        case FK_2D_VEL_XN: zh2_vel_xn<T>(f, ux, r0); break;
        case FK_2D_VEL_XP: zh2_vel_xp<T>(f, ux, r0); break;
        case FK_2D_DENS_XP: zh2_dens_xp<T>(f, T(F.rho), r0); break;
        case FK_2D_VEL_YN: zh2_vel_yn<T>(f, ux, r0); break;
        case FK_2D_VEL_YP: zh2_vel_yp<T>(f, ux, r0); break;
        default: break;
      }
    } else {
      switch (F.kind) {
        case FK_3D_VEL_XN: zh3_vel_xn<T>(f, ux, uy, uz, r0); break;
        case FK_3D_VEL_XP: zh3_vel_xp<T>(f, ux, uy, uz, r0); break;
        case FK_3D_DENS_XP: zh3_dens_xp<T>(f, T(F.rho), uy, uz, r0); break;
        case FK_3D_VEL_YN: zh3_vel_yn<T>(f, ux, uy, uz, r0); break;
        case FK_3D_VEL_YP: zh3_vel_yp<T>(f, ux, uy, uz, r0); break;
        case FK_3D_VEL_ZN: zh3_vel_zn<T>(f, ux, uy, uz, r0); break;
        case FK_3D_VEL_ZP: zh3_vel_zp<T>(f, ux, uy, uz, r0); break;
        default: break;
      }
    }
  }
}

}  // namespace lbmpsm
