// Internal interfaces between the C ABI layer (api.hip) and the kernel
// translation units.  Nothing here is exported.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lbm_common.h"

namespace lbmpsm {

// Faces of one `lbm-psm-bc` keyword (at most 6), or of its emulation on a
// reference rank grid (one segment per rank along each tangential axis).
constexpr int kMaxFaces = 256;

// Stored lattice geometry.  Node (i,j,k) of the stored box sits at
// n = (i*sy + j)*sz + k; each population is a plane of `plane` elements.
// Periodic axes hold exactly the global nodes and wrap in the kernels; a
// non-periodic axis keeps the reference's halo layer on both sides, so its
// stored coordinate equals the reference's halo'd index.  `off` is the
// reference index minus the stored index per axis.
struct Geom {
  int dim, q;
  int sx, sy, sz;
  int per[3];
  int off[3];
  int nx, ny, nz;  // global interior counts (nz = 1 in 2D)
  int64_t nodes;   // sx*sy*sz
  int64_t plane;   // nodes rounded up to a multiple of 64
};

// Index of direction q at stored node n in an FP64 population buffer of Q
// directions and `plane` elements per direction: blocks of 64 nodes that hold
// their Q directions as 64-element runs (LBMPSM_F64_BLOCKED = 1), or one plane
// per direction (0).  The collide, the read-back kernels and the adjoint's
// reads of recomputed states use it.
#ifndef LBMPSM_F64_BLOCKED
#define LBMPSM_F64_BLOCKED 1
#endif
template <int Q>
__host__ __device__ inline size_t pop_index_f64(int q, int64_t plane, unsigned n) {
  // the node's part first, so that the directions of one node share it
  if constexpr (LBMPSM_F64_BLOCKED)
    return ((size_t)(n >> 6) * (Q * 64) + (n & 63u)) + (size_t)q * 64;
  else return (size_t)q * (size_t)plane + n;
}

// Pull-stream + BC + collide for one node.  fin/fout point at plane 0 of a
// population buffer (element type depends on the storage form).
struct CollideArgs {
  Geom g;
  const void* fin;
  void* fout;
  uint8_t* mask;        // kMaskOcc*: slot holds a particle; kMaskFresh*: set
                        // this step (force starts at zero; cleared here)
  const double* sf;     // [2][plane]
  const double* us;     // [6][plane]   (3*s + d)
  double* Fh;           // [6][plane]   (3*s + d), lattice units
  const float* sff;     // sf, us, Fh in float (FP32 / FP16 storage)
  const float* usf;
  float* Fhf;
  double* rho_out;      // [plane]      or null
  double* u_out;        // [3][plane]   or null
  int poststream;       // 1: fin already holds post-stream values (no pull, no BC)
  int guo;              // |F|^2 > 0
  double omega;
  double guo_pref;      // 1 - 0.5*omega
  double F[3];
  double invCs2, invCs4;
  double hInvCs4;       // 0.5 * invCs4 (exact)
  double rho0;          // reference density of shifted storage
  // the FP32 form's per-launch constants, rounded once on the host: omega,
  // rho0, F[3], 3 w_axis F[3] / 2 (the half force of the odd equilibrium),
  // 1 - omega/2
  float lowp_c[9];
  float scale, inv_scale;  // FP16 storage scaling (power of two)
  const double* cb;     // [q] value of populations pulled from beyond a non-periodic
                        // halo, in storage form (device memory)
  const float* cbf;     // the same rounded to float, for the FP32 form
  const Face* faces;    // [nfaces] Zou-He faces in application order (device memory)
  int nfaces;
};

// Workgroups per entry of the force/torque/stresslet reduction.
#ifndef LBMPSM_HYDRO_GROUPS
#define LBMPSM_HYDRO_GROUPS 2
#endif
constexpr int kHydroGroups = LBMPSM_HYDRO_GROUPS;

// Per-node slot mask bits: slot s holds a particle (tag > 0); slot s was set
// by this step's map, so its node force starts from zero in the collide
// (which then clears the bit).
constexpr uint8_t kMaskOcc0 = 1, kMaskOcc1 = 2, kMaskFresh0 = 4, kMaskFresh1 = 8;

// Particle list state on the device.  The raw LAMMPS arrays of one call
// arrive in one copy: x[n][3], v[n][3], omega[n][3], radius[n] (doubles)
// followed by tag[n] (ints), i.e. 84 bytes per entry.
struct ParticleDev {
  int cap = 0;
  int n = 0;
  char* raw = nullptr;      // [cap * 84] bytes
  // derived
  double* xl = nullptr;     // [n][4]  x_loc, y_loc, z_loc (reference halo'd coords), r_lb
  double* kin = nullptr;    // [n][6]  uTrans, omegaLB
  int* zone = nullptr;      // [n][6]  stored-coordinate map zone [lo,hi) x,y,z
  int* rzone = nullptr;     // [n][6]  stored-coordinate reduction zone
  // the same arrays of the previous list: a new list is prepared in them
  // while the step of the previous one may still read its own
  char* raw_alt = nullptr;
  double* xl_alt = nullptr;
  double* kin_alt = nullptr;
  int* zone_alt = nullptr;
  int* rzone_alt = nullptr;
  double* out = nullptr;    // [n][12] F, T, S (physical units)
  double* hpart = nullptr;  // [n][kHydroGroups][12] group sums of the reduction
  const double* x(int nn) const { return (const double*)raw; }
  const double* v(int nn) const { return (const double*)raw + 3 * (size_t)nn; }
  const double* w(int nn) const { return (const double*)raw + 6 * (size_t)nn; }
  const double* r(int nn) const { return (const double*)raw + 9 * (size_t)nn; }
  const int* tag(int nn) const { return (const int*)(raw + 80 * (size_t)nn); }
};

// Node bricks of the particle map: per brick, the list entries whose map
// zone touches it (unordered, at most kBrickCap; longer lists make the map
// kernel scan the whole particle list for that brick).
constexpr int kBrickCap = 64;
// Brick edge lengths (x, y, z nodes; 256 nodes, one per thread of a map
// workgroup): 3D, and 2D (z extent 1).
#ifndef LBMPSM_BRICK3_I
#define LBMPSM_BRICK3_I 4
#endif
#ifndef LBMPSM_BRICK3_J
#define LBMPSM_BRICK3_J 4
#endif
constexpr int kBrick3[3] = {LBMPSM_BRICK3_I, LBMPSM_BRICK3_J,
                            256 / (LBMPSM_BRICK3_I * LBMPSM_BRICK3_J)};
constexpr int kBrick2[3] = {16, 16, 1};
static_assert(kBrick3[0] * kBrick3[1] * kBrick3[2] == 256, "256-node bricks");
struct BrickDev {
  int bi = 0, bj = 0, bk = 0;  // brick edge lengths (nodes)
  int nbi = 0, nbj = 0, nbk = 0;
  int64_t nbricks = 0;
  int* count = nullptr;     // [nbricks]
  int* list = nullptr;      // [nbricks][kBrickCap]
  int* slow_list = nullptr; // [nbricks] bricks for the general map kernel
  int* slow_count = nullptr;  // [1]
};

struct SlotsDev {
  int* tag = nullptr;       // [2][plane]
  double* sf = nullptr;     // [2][plane]
  double* us = nullptr;     // [6][plane]
  double* Fh = nullptr;     // [6][plane]
  uint8_t* mask = nullptr;  // [plane]
  // FP32 / FP16 storage: solid fractions, node velocities and node forces
  // are kept in float, in the first half of sf / us / Fh (null for FP64
  // storage)
  float* sff = nullptr;
  float* usf = nullptr;
  float* Fhf = nullptr;
};

// Solid-fraction code of an entry at a node of its map zone: the
// sub-point count (fraction * count), or kSfCodeOne for the reference's
// early "fully inside" exit (exactly 1.0).
constexpr unsigned short kSfCodeOne = 0xFFFF;

struct MapArgs {
  Geom g;
  SlotsDev s;
  int n;
  const int* tag;
  const double* xl;
  const double* kin;
  const int* zone;
  const int* list;
  const int* count;
  int* count_reset;        // the same counts, zeroed by k_map once read
  unsigned short* sfcode;  // [n][zone_cap] codes of each entry's zone, k fastest
  int zone_cap;            // codes per entry
  int zone_span;           // bound of every zone's extent per axis
  int* slow_list;          // [nbricks] bricks k_map leaves to k_map_general
  int* slow_count;         // their number (zeroed by k_prep_bin)
  int bi, bj, bk, nbi, nbj, nbk;
  double hsq_half;
};

// ---- launchers (lattice_f64.hip / lattice_lowp.hip) ------------------------
// Bulk nodes (every node on no Zou-He face) ...
hipError_t launch_collide(int storage, const CollideArgs& a, hipStream_t s);
// ... and the face nodes, with the faces applied after the pull.
hipError_t launch_fill_planes(int storage, void* f, const Geom& g, const double* cb,
                              double rho0, float scale, hipStream_t s);
// Post-stream + BC values of interior nodes into `out` (FP64, reference
// order [nx][ny][nz][q]).
hipError_t launch_finish(int storage, const CollideArgs& a, double* out, hipStream_t s);
// Interior reference-order FP64 populations (post-stream) into storage form.
hipError_t launch_set_populations(int storage, void* f, const Geom& g, const double* in,
                                  double rho0, float scale, hipStream_t s);
// The same in the reference's full halo'd layout ([nx+2][ny+2][nz+2 or 1][q]):
// periodic halo positions show their periodic image on output and are
// skipped on input.
hipError_t launch_finish_ref(int storage, const CollideArgs& a, double* out, hipStream_t s);
hipError_t launch_set_populations_ref(int storage, void* f, const Geom& g, const double* in,
                                      double rho0, float scale, hipStream_t s);
// Mass of the post-stream state: per-block partial sums into `partials`.
hipError_t launch_mass(int storage, const CollideArgs& a, double* partials, int nblocks,
                       hipStream_t s);

// ---- particles.hip ----------------------------------------------------------
struct PrepArgs {
  Geom g;
  int n;
  const double* x;
  const double* v;
  const double* w;
  const double* r;
  const int* tag_in;  // the list's tags
  int* tag_out;       // where the later kernels read them (null: tag_in is that copy)
  double* xl;
  double* kin;
  int* zone;
  int* rzone;
  double origin[3];
  double dx, Uc, u_lb, tc, dt_d;
  int* count;
  int* list;
  int* slow_count;  // work-list length of k_map_general, zeroed here
  int bi, bj, bk, nbi, nbj, nbk;
};
hipError_t launch_fill_int(int* p, int64_t n, int value, hipStream_t s);
// Per-entry lattice quantities and brick lists (one wavefront per entry).
hipError_t launch_prep_bin(const PrepArgs& a, hipStream_t s);
// Solid fraction of every entry at every node of its map zone.
hipError_t launch_entry_sf(const MapArgs& a, hipStream_t s);
// Slot state machine on every touched brick.
hipError_t launch_map(const MapArgs& a, int64_t nbricks, hipStream_t s);
struct HydroArgs {
  Geom g;
  int n;
  const int* tag;
  const double* xl;
  const int* rzone;
  const int* slot_tag;
  const uint8_t* mask;  // slot occupancy: a slot without its bit holds tag 0
  const double* Fh;
  const float* Fhf;  // node forces in float (FP32 / FP16 storage), else null
  double forceFactor, torqueFactor;
  double* part;  // [n][kHydroGroups][12] group sums
  double* out;
};
hipError_t launch_hydro(const HydroArgs& a, hipStream_t s);
hipError_t launch_slots_export(const Geom& g, const SlotsDev& sl, int* tag, double* sf,
                               double* us, double* F, hipStream_t s);
hipError_t launch_macros_export(const Geom& g, const double* rho, const double* u,
                                double* rho_out, double* u_out, hipStream_t s);
hipError_t launch_slots_reset(const Geom& g, const SlotsDev& sl, hipStream_t s);

// ---- adjoint.hip -------------------------------------------------------------
// One reverse step of the FP64 step on a fully periodic lattice without
// faces: lambda^t = J^T (pull^T lambda^{t+1}) plus the node-force seed on the
// last step; parameter gradients accumulate into acc[10][plane]
// (om, F[3], u_s slot 0 [3], u_s slot 1 [3]).
struct AdjArgs {
  Geom g;
  const double* hprev;    // [q][plane] state before the step (post-collision of t-1)
  int inplace;            // 1: hprev holds post-stream values (read at the node)
  const double* lam_in;   // [q][plane] adjoint of the next step's input, or null
  double* lam_out;        // [q][plane]
  const uint8_t* mask;
  const double* sf;       // [2][plane]
  const double* us;       // [6][plane]
  const double* seed;     // [6][plane] d loss / d node force of this step, or null
  double* acc;            // [10][plane]
  double omega, invCs2, invCs4;
  double F[3];
};
struct AdjSeedArgs {
  Geom g;
  int n;
  const int* tag;
  const double* xl;
  const int* zone;
  const int* rzone;
  const int* slot_tag;   // [2][plane]
  const uint8_t* mask;
  const double* bar;     // [n][12] d loss / d (F, T, S) per entry
  double* seed;          // [6][plane]
  int* writer;           // [2][plane]
  double forceFactor, torqueFactor;
};
hipError_t launch_adj_step(const AdjArgs& a, hipStream_t s);
hipError_t launch_adj_seed(const AdjSeedArgs& a, hipStream_t s);
hipError_t launch_adj_writer(const AdjSeedArgs& a, hipStream_t s);
// out[n][6]: d/dv_lb, d/dw_lb per entry
hipError_t launch_adj_entry(const AdjSeedArgs& a, const double* acc, double* out, hipStream_t s);
// partials[nblocks][4]: per-block sums of acc planes 0..3
hipError_t launch_adj_param(const Geom& g, const double* acc, double* partials, int nblocks,
                            hipStream_t s);

}  // namespace lbmpsm
