// Lattice kernels: fused pull-stream + Zou-He + BGK-PSM collide, and the
// helpers that read the lattice back in the reference's form.
//
// State between steps is the post-collision population of every stored node
// (blocks of 64 nodes, each holding its directions as 64-element runs;
// see pop_index).  A step pulls
// f_q(x) = S_q(x - e_q), applies the Zou-He faces of the previous step to
// the pulled node, computes rho/u, collides and writes S'(x) to the other
// buffer.  This is the reference's order (collide, push-stream, BC at the
// destination; src/LBM_PSM_BGK_dynamics.cpp:81-109, src/fix_LBM_PSM_BC.cpp)
// regrouped per node so that every population is read once and written once.
//
// Storage forms: F64 keeps raw populations and the reference's FP64
// arithmetic (this TU is compiled with -ffp-contract=off).  F32 and F16 keep
// g = f - w*rho0 (F16 scaled by a power of two) and compute in FP32 with
// contraction enabled locally.
#include <hip/hip_fp16.h>

#include <algorithm>
#include <type_traits>

#include "lbmpsm_internal.h"

// Tuning knobs of the streaming kernels (compile-time).
#ifndef LBMPSM_NT_STORE
#define LBMPSM_NT_STORE 1  // non-temporal population stores
#endif
#ifndef LBMPSM_NT_LOAD
#define LBMPSM_NT_LOAD 1   // non-temporal population loads
#endif
#ifndef LBMPSM_F16_DWORD_LOAD
#define LBMPSM_F16_DWORD_LOAD 0  // FP16: load the dword holding the element
#endif
#ifndef LBMPSM_WAVES_PER_EU
#define LBMPSM_WAVES_PER_EU 0  // > 0: ask for this many waves per SIMD
#endif
#if LBMPSM_WAVES_PER_EU > 0
#define LBMPSM_OCC __attribute__((amdgpu_waves_per_eu(LBMPSM_WAVES_PER_EU)))
#else
#define LBMPSM_OCC
#endif
#ifndef LBMPSM_F64_SPLIT
#define LBMPSM_F64_SPLIT 1  // 1: FP64 particle waves without a second particle skip its terms
#endif
#ifndef LBMPSM_LOWP_WAVES
#define LBMPSM_LOWP_WAVES 4  // > 0: waves per SIMD asked of the bulk k_collide_lowp
#endif
#if LBMPSM_LOWP_WAVES > 0
// the once-per-run post-stream variant is left to the compiler (it would spill)
#define LBMPSM_LOWP_OCC __attribute__((amdgpu_waves_per_eu(POST ? 1 : LBMPSM_LOWP_WAVES)))
#else
#define LBMPSM_LOWP_OCC
#endif
#ifndef LBMPSM_ROW_WAVES
#define LBMPSM_ROW_WAVES 1  // waves per bulk workgroup; 0: the chunks of one row, at most 4
#endif
#ifndef LBMPSM_ROW_WAVES_LOWP
#define LBMPSM_ROW_WAVES_LOWP 2  // the same for the FP32 / FP16 kernels on a large lattice
#endif
#ifndef LBMPSM_LARGE_WAVES
#define LBMPSM_LARGE_WAVES 262144  // a large lattice: at least this many 64-node chunks
#endif
static_assert(LBMPSM_ROW_WAVES >= 0 && LBMPSM_ROW_WAVES <= 4, "bulk workgroups hold at most 4 waves");
static_assert(LBMPSM_ROW_WAVES_LOWP >= 0 && LBMPSM_ROW_WAVES_LOWP <= 4,
              "bulk workgroups hold at most 4 waves");
#ifndef LBMPSM_XCD_SWIZZLE
#define LBMPSM_XCD_SWIZZLE 1  // 1: give each XCD a contiguous range of workgroups
#endif
#ifndef LBMPSM_XCD_SWIZZLE_MIN
#define LBMPSM_XCD_SWIZZLE_MIN 131072  // ... in grids of at least this many workgroups
#endif

namespace lbmpsm {

enum { ST_F64 = 0, ST_F32 = 1, ST_F16 = 2 };

template <int ST>
struct StoreT;
template <>
struct StoreT<ST_F64> {
  using T = double;
  using C = double;
};
template <>
struct StoreT<ST_F32> {
  using T = float;
  using C = float;
};
template <>
struct StoreT<ST_F16> {
  using T = __half;
  using C = float;
};

// NT: a non-temporal load (the line is not kept for reuse).
template <int ST, bool NT = (LBMPSM_NT_LOAD != 0)>
__device__ __forceinline__ typename StoreT<ST>::C ld(const typename StoreT<ST>::T* p,
                                                      float inv_scale) {
  if constexpr (ST == ST_F16) {
#if LBMPSM_F16_DWORD_LOAD
    // the aligned dword that holds the element (neighbouring lanes load the
    // same dword), then its half
    const size_t ad = (size_t)p;
    const unsigned w = *(const unsigned*)(ad & ~(size_t)3);
    const unsigned short b = (unsigned short)((ad & 2) ? (w >> 16) : (w & 0xffffu));
    return __half2float(__ushort_as_half(b)) * inv_scale;
#else
    if constexpr (NT) {
      const unsigned short b = __builtin_nontemporal_load((const unsigned short*)p);
      return __half2float(__ushort_as_half(b)) * inv_scale;
    } else {
      return __half2float(*p) * inv_scale;
    }
#endif
  } else {
    (void)inv_scale;
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
  }
}

template <int ST>
__device__ __forceinline__ void st(typename StoreT<ST>::T* p, typename StoreT<ST>::C v,
                                   float scale) {
  if constexpr (ST == ST_F16) {
#if LBMPSM_NT_STORE
    __builtin_nontemporal_store(__half_as_ushort(__float2half_rn(v * scale)), (unsigned short*)p);
#else
    *p = __float2half_rn(v * scale);
#endif
  } else {
    (void)scale;
#if LBMPSM_NT_STORE
    __builtin_nontemporal_store(v, p);
#else
    *p = v;
#endif
  }
}

// Neighbour coordinates c-1 and c+1 along an axis of n nodes.  Off a
// non-periodic edge the flag is cleared and the coordinate is left at c, so
// every address formed from it stays inside the buffer.
__device__ __forceinline__ void nbr(int c, int n, int per, int& cm, bool& vm, int& cp,
                                    bool& vp) {
  cm = c - 1;
  vm = true;
  cp = c + 1;
  vp = true;
  if (cm < 0) {
    if (per) {
      cm += n;
    } else {
      vm = false;
      cm = c;
    }
  }
  if (cp >= n) {
    if (per) {
      cp -= n;
    } else {
      vp = false;
      cp = c;
    }
  }
}

// One wavefront streams 64 consecutive nodes of one row of the stored box:
// a z-row (i,j) in 3D, a y-row i in 2D.  Everything that depends only on the
// row is wave-uniform (scalar registers); a lane carries its column and
// 32-bit byte offsets, so every access is a uniform 64-bit base plus a
// 32-bit lane offset.
struct Row {
  int i, j;        // row coordinates (j unused in 2D)
  int im, ip, jm, jp;
  unsigned okmask; // bit 3a+b: neighbour row (i+a-1, j+b-1) exists
  int ncols;
  int c0;          // first column of this wave
};

template <int DIM>
__device__ __forceinline__ bool make_row(const Geom& g, int64_t wave64, Row& r) {
  // 32-bit arithmetic: the stored box has fewer than 2^31 nodes
  bool vim, vip, vjm = true, vjp = true;
  const int wave = (int)wave64;
  const int ncols = DIM == 3 ? g.sz : g.sy;
  const int nch = (ncols + 63) >> 6;
  const int row = wave / nch;
  r.c0 = (wave - row * nch) * 64;
  r.ncols = ncols;
  if constexpr (DIM == 3) {
    if (row >= g.sx * g.sy) return false;
    r.i = row / g.sy;
    r.j = row - r.i * g.sy;
    nbr(r.i, g.sx, g.per[0], r.im, vim, r.ip, vip);
    nbr(r.j, g.sy, g.per[1], r.jm, vjm, r.jp, vjp);
  } else {
    if (row >= g.sx) return false;
    r.i = row;
    r.j = 0;
    r.jm = r.jp = 0;
    nbr(r.i, g.sx, g.per[0], r.im, vim, r.ip, vip);
  }
  const bool iv[3] = {vim, true, vip}, jv[3] = {vjm, true, vjp};
  unsigned m = 0;
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b)
      if (iv[a] && jv[b]) m |= 1u << (3 * a + b);
  r.okmask = m;
  return true;
}

// Node index of column 0 of the row at offset (a-1, b-1) from row r.
template <int DIM>
__device__ __forceinline__ unsigned row_base(const Geom& g, const Row& r, int a, int b) {
  const int ii = a == 0 ? r.im : (a == 2 ? r.ip : r.i);
  if constexpr (DIM == 3) {
    const int jj = b == 0 ? r.jm : (b == 2 ? r.jp : r.j);
    return (unsigned)((ii * g.sy + jj) * g.sz);
  } else {
    return (unsigned)(ii * g.sy);
  }
}

// Planar layout (one plane of `plane` elements per direction), kept for the
// LBMPSM_*_BLOCKED = 0 builds: uniform base, 32-bit lane offset.
template <typename T>
__device__ __forceinline__ const T* pq_ptr(const void* base, int q, int64_t plane, unsigned rowb) {
  return (const T*)((const char*)base + ((size_t)q * (size_t)plane + rowb) * sizeof(T));
}
template <typename T>
__device__ __forceinline__ T* pq_ptr_w(void* base, int q, int64_t plane, unsigned rowb) {
  return (T*)((char*)base + ((size_t)q * (size_t)plane + rowb) * sizeof(T));
}
template <typename T>
__device__ __forceinline__ T* lane_at(T* p, unsigned col) {
  return (T*)((char*)p + col * (unsigned)sizeof(T));
}

// Storage index of direction q at stored node n: blocks of 64 nodes that
// hold their Q directions as 64-element runs, so that the loads and stores
// of a wave touch one contiguous region per node row instead of one per
// direction (FP64: pop_index_f64, which the adjoint shares).
#ifndef LBMPSM_LOWP_BLOCKED
#define LBMPSM_LOWP_BLOCKED 1
#endif
template <int ST, int Q>
__device__ __forceinline__ size_t pop_index(int q, int64_t plane, unsigned n) {
  if constexpr (ST == ST_F64) return pop_index_f64<Q>(q, plane, n);
  else if constexpr (!LBMPSM_LOWP_BLOCKED) return (size_t)q * (size_t)plane + n;
  else return ((size_t)(n >> 6) * (Q * 64) + (n & 63u)) + (size_t)q * 64;
}
template <int ST>
constexpr bool kPlanar = ST == ST_F64 ? !LBMPSM_F64_BLOCKED : !LBMPSM_LOWP_BLOCKED;
// Element of direction q at column col of the row starting at node rowb.
template <int ST, int Q>
__device__ __forceinline__ const typename StoreT<ST>::T* pop_ptr(const void* base, int q,
                                                                 int64_t plane, unsigned rowb,
                                                                 unsigned col) {
  using T = typename StoreT<ST>::T;
  if constexpr (kPlanar<ST>)
    return lane_at(pq_ptr<T>(base, q, plane, rowb), col);
  else
    return (const T*)base + pop_index<ST, Q>(q, plane, rowb + col);
}
template <int ST, int Q>
__device__ __forceinline__ typename StoreT<ST>::T* pop_ptr_w(void* base, int q, int64_t plane,
                                                             unsigned rowb, unsigned col) {
  using T = typename StoreT<ST>::T;
  if constexpr (kPlanar<ST>)
    return lane_at(pq_ptr_w<T>(base, q, plane, rowb), col);
  else
    return (T*)base + pop_index<ST, Q>(q, plane, rowb + col);
}

// Rows whose length is a multiple of 64 (ALN) start on a block boundary, so
// the blocked index of (row start rowb, column col) splits into a uniform
// part, the row's first block plus the direction's run, and a lane part
// that depends on the column only: one 32-bit lane offset then serves
// every direction pulled from the same column, and each access is a scalar
// base plus that offset.
template <int ST, int Q>
__device__ __forceinline__ unsigned col_off(unsigned col) {
  return ((col >> 6) * (unsigned)(Q * 64) + (col & 63u)) * (unsigned)sizeof(typename StoreT<ST>::T);
}
// The uniform part as a scalar 64-bit base: passing it through readfirstlane
// keeps the compiler from folding the direction's run into the lane offset
// (a 64-bit vector add per access when the run lies beyond the 4 KiB
// immediate offset of a global access).
// The result is marked as global memory (address space 1), which an integer
// round trip would otherwise lose (flat accesses).
__device__ __forceinline__ char* scalar_base(const void* base, size_t u) {
  using gchar = __attribute__((address_space(1))) char;
  const uint64_t b = (uint64_t)base + u;
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)b);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(b >> 32));
  return (char*)(gchar*)(((uint64_t)hi << 32) | lo);
}
// Direction q's run starts q * 64 elements into the block; the part of that
// below 4 KiB is left to the access's immediate offset, so the directions
// of one row share a few scalar bases.
template <int ST, int q>
constexpr int kRunHi = (q * 64 * (int)sizeof(typename StoreT<ST>::T)) & ~4095;
template <int ST, int q>
constexpr int kRunLo = (q * 64 * (int)sizeof(typename StoreT<ST>::T)) & 4095;
template <int ST, int Q, int q>
__device__ __forceinline__ const typename StoreT<ST>::T* pop_ptr_al(const void* base,
                                                                    unsigned rowb, unsigned off) {
  using T = typename StoreT<ST>::T;
  const size_t u = (size_t)(rowb >> 6) * (Q * 64 * sizeof(T)) + kRunHi<ST, q>;
  return (const T*)(scalar_base(base, u) + off + kRunLo<ST, q>);
}
template <int ST, int Q, int q>
__device__ __forceinline__ typename StoreT<ST>::T* pop_ptr_al_w(void* base, unsigned rowb,
                                                               unsigned off) {
  using T = typename StoreT<ST>::T;
  const size_t u = (size_t)(rowb >> 6) * (Q * 64 * sizeof(T)) + kRunHi<ST, q>;
  return (T*)(scalar_base(base, u) + off + kRunLo<ST, q>);
}

// Pull the post-stream populations of the node in column c of row r, then
// apply the Zou-He faces.  In post-stream mode the node's own values are
// read and no face is applied (initial state / state set by the host).
// PS: 0 pull, 1 post-stream, 2 as a.poststream says (the bulk kernels are
// compiled for one mode, which keeps the other path's loads out of their
// wait-count bookkeeping).  ALN: rows are a whole number of blocks long
// (col_off).
template <int DIM, int Q, int ST, bool FACES, int PS = 2, bool ALN = false>
__device__ __forceinline__ void gather_row(const CollideArgs& a, const Row& r, int c,
                                           typename StoreT<ST>::C* f) {
  using T = typename StoreT<ST>::T;
  using C = typename StoreT<ST>::C;
  constexpr bool AL = ALN && !kPlanar<ST>;
  const Geom& g = a.g;
  const int64_t plane = g.plane;
  if (PS == 1 || (PS == 2 && a.poststream)) {
    const unsigned rb = row_base<DIM>(g, r, 1, 1);
    Unroll<0, Q>::run([&]<int q>() {
      f[q] = ld<ST>(pop_ptr<ST, Q>(a.fin, q, plane, rb, (unsigned)c), a.inv_scale);
    });
    return;
  }
  const int per_c = DIM == 3 ? g.per[2] : g.per[1];
  int cm, cp;
  bool vcm, vcp;
  nbr(c, r.ncols, per_c, cm, vcm, cp, vcp);
  unsigned oc = 0, om = 0, op = 0;
  if constexpr (AL) {
    oc = col_off<ST, Q>((unsigned)c);
    om = col_off<ST, Q>((unsigned)cm);
    op = col_off<ST, Q>((unsigned)cp);
  }
  // every load unconditionally (src stays inside the buffer: nbr keeps the
  // coordinate off a non-periodic edge), so that the loads of all
  // directions are in flight together
  Unroll<0, Q>::run([&]<int q>() {
    constexpr int ex = VelSet<Q>::ex[q], ey = VelSet<Q>::ey[q], ez = VelSet<Q>::ez[q];
    // source node = node - e_q; the column axis is z in 3D and y in 2D
    constexpr int ea = ex;
    constexpr int eb = DIM == 3 ? ey : 0;
    constexpr int ec = DIM == 3 ? ez : ey;
    const unsigned col = (unsigned)(ec == 1 ? cm : (ec == -1 ? cp : c));
    const unsigned rs = row_base<DIM>(g, r, 1 - ea, 1 - eb);
    const T* src;
    if constexpr (AL) src = pop_ptr_al<ST, Q, q>(a.fin, rs, ec == 1 ? om : (ec == -1 ? op : oc));
    else src = pop_ptr<ST, Q>(a.fin, q, plane, rs, col);
    f[q] = ld<ST>(src, a.inv_scale);
  });
  // then the wall default for a source off a non-periodic edge; a wave whose
  // neighbour rows all exist, on a periodic column axis, has none and skips
  // the selects (with the default in a scalar register, each also costs a
  // move into a vector register)
  if (r.okmask != 0x1FFu || !per_c) {
    Unroll<0, Q>::run([&]<int q>() {
      constexpr int ea = VelSet<Q>::ex[q];
      constexpr int eb = DIM == 3 ? VelSet<Q>::ey[q] : 0;
      constexpr int ec = DIM == 3 ? VelSet<Q>::ez[q] : VelSet<Q>::ey[q];
      const bool cok = ec == 1 ? vcm : (ec == -1 ? vcp : true);
      const bool rok = (r.okmask >> (3 * (1 - ea) + (1 - eb))) & 1u;
      C cbv;
      if constexpr (ST == ST_F64) cbv = a.cb[q];
      else cbv = a.cbf[q];
      f[q] = (rok && cok) ? f[q] : cbv;
    });
  }
  if constexpr (FACES) {
    const C r0 = (ST == ST_F64) ? C(0) : (C)a.rho0;
    const int jj = DIM == 3 ? r.j : c;
    const int kk = DIM == 3 ? c : 0;
    apply_faces<Q, C>(f, a.faces, a.nfaces, r.i, jj, kk, r0);
  }
}

// True when node (i,j,k) lies on one of the faces (plane and ranges).
__device__ __forceinline__ bool on_face(const Face* faces, int nfaces, int i, int j, int k) {
  for (int fi = 0; fi < nfaces; ++fi) {
    const Face& F = faces[fi];
    const int c = F.axis == 0 ? i : (F.axis == 1 ? j : k);
    if (c != F.plane) continue;
    const int aa = F.axis == 0 ? j : i;
    const int bb = F.axis == 2 ? j : k;
    if (aa < F.a_lo || aa > F.a_hi || bb < F.b_lo || bb > F.b_hi) continue;
    return true;
  }
  return false;
}

// ---------------------------------------------------------------------------
// FP64 collision, reference expression order
// (src/LBM_PSM_BGK_dynamics.h:40-114, src/LBM_PSM_dynamics.h:43-116).
// ---------------------------------------------------------------------------
// Bulk kernels run one wave per 64-node chunk of a row; a workgroup holds
// several consecutive waves, normally the chunks of one row, so that the
// cache lines a z-shifted population straddles between chunks are fetched
// once per compute unit.
__device__ __forceinline__ int bulk_wave() {
  int b = (int)blockIdx.x;
#if LBMPSM_XCD_SWIZZLE
  // workgroups are dealt round-robin over the 8 XCDs; renumber them so
  // that each XCD (and its L2) walks one contiguous range of rows, whose
  // neighbouring rows it then reads from its own L2 (a large lattice; a
  // small one is served from the shared last-level cache anyway)
  const int nb = (int)gridDim.x;
  if (nb >= LBMPSM_XCD_SWIZZLE_MIN) {
    const int per = nb >> 3, rem = nb & 7, x = b & 7;
    b = x * per + min(x, rem) + (b >> 3);
  }
#endif
  // wave-uniform by construction; readfirstlane lets the compiler keep the
  // row set-up in scalar registers and address with scalar bases
  return __builtin_amdgcn_readfirstlane(b * (int)(blockDim.x >> 6) + (int)(threadIdx.x >> 6));
}

// Slot data of a node for the FP64 particle form: B = solid fraction, u_s =
// rigid-body node velocity, Fa = node force carried over (zero for a slot
// the map set this step).  psm: some lane of the wave holds a particle.
struct SlotsF64 {
  uint8_t m;
  bool psm, o1, o2;
  double B1, B2;
  double us1[3], us2[3];
  double Fa[6];
};

__device__ __forceinline__ void load_slots_f64(const CollideArgs& a, unsigned n, uint8_t m,
                                               SlotsF64& S) {
  const int64_t plane = a.g.plane;
  S.m = m;
  S.psm = __any(m != 0);
  S.o1 = m & 1;
  S.o2 = m & 2;
  S.B1 = S.B2 = 0.0;
  for (int d = 0; d < 3; ++d) S.us1[d] = S.us2[d] = 0.0;
  for (int d = 0; d < 6; ++d) S.Fa[d] = 0.0;
  if (!S.psm) return;
  if (S.o1) {
    S.B1 = a.sf[n];
    S.us1[0] = a.us[n];
    S.us1[1] = a.us[plane + n];
    S.us1[2] = a.us[2 * plane + n];
    if (!(m & kMaskFresh0)) {
      S.Fa[0] = a.Fh[n];
      S.Fa[1] = a.Fh[plane + n];
      S.Fa[2] = a.Fh[2 * plane + n];
    }
  }
  if (S.o2) {
    S.B2 = a.sf[plane + n];
    S.us2[0] = a.us[3 * plane + n];
    S.us2[1] = a.us[4 * plane + n];
    S.us2[2] = a.us[5 * plane + n];
    if (!(m & kMaskFresh1)) {
      S.Fa[3] = a.Fh[3 * plane + n];
      S.Fa[4] = a.Fh[4 * plane + n];
      S.Fa[5] = a.Fh[5 * plane + n];
    }
  }
}

// Collision of one node whose post-stream populations are in f[] (FP64).
template <int DIM, int Q, int FAX, bool ALN>
__device__ __forceinline__ void collide_f64(const CollideArgs& a, double* f, unsigned rb,
                                            unsigned uc, SlotsF64& S);

// Guo's forcing term of direction q without its prefactor (1 - omega/2),
// src/LBM_PSM_BGK_dynamics.h:40-114: w ((e - u).F / cs^2 + (e.u)(e.F) / cs^4),
// in the reference's operand order.  FAX >= 0: only F[FAX] is nonzero (a
// body force along one lattice axis, as gravity usually is); the terms of
// the other components are then exact zeros, and leaving them out of the
// sums changes a result at most in the sign of a zero.
template <int Q, int q, int FAX>
__device__ __forceinline__ double guo_f64(double ux, double uy, double uz, double eu, double F0,
                                          double F1, double F2, double inv2, double inv4) {
  constexpr double w = VelSet<Q>::w[q];
  constexpr int cx = VelSet<Q>::ex[q], cy = VelSet<Q>::ey[q], cz = VelSet<Q>::ez[q];
  constexpr double ex = cx, ey = cy, ez = cz;
  if constexpr (FAX < 0) {
    const double eF = edot<Q, q>(F0, F1, F2);
    return w * (((ex - ux) * F0 + (ey - uy) * F1 + (ez - uz) * F2) * inv2 + (eu * inv4 * eF));
  } else {
    constexpr int ca = FAX == 0 ? cx : (FAX == 1 ? cy : cz);
    constexpr double ea = ca;
    const double Fa = FAX == 0 ? F0 : (FAX == 1 ? F1 : F2);
    const double ua = FAX == 0 ? ux : (FAX == 1 ? uy : uz);
    if constexpr (ca != 0) return w * ((ea - ua) * Fa * inv2 + (eu * inv4 * eterm<ca>(Fa)));
    else return w * ((ea - ua) * Fa * inv2);
  }
}

// Bulk: every stored node, one wave per 64-node chunk of a row.  A node on
// Zou-He faces has them applied, in table order, to its pulled populations;
// its pulls are part of the row's coalesced loads (a z face node would
// otherwise be a strided access of its own).
template <int DIM, int Q, bool POST, int FAX, bool ALN>
__global__ __launch_bounds__(256) LBMPSM_OCC void k_collide_f64(const CollideArgs a) {
  const Geom& g = a.g;
  Row r;
  if (!make_row<DIM>(g, bulk_wave(), r)) return;
  const int c = r.c0 + (int)(threadIdx.x & 63);
  if (c >= r.ncols) return;
  const bool face = !POST && a.nfaces &&
      on_face(a.faces, a.nfaces, r.i, DIM == 3 ? r.j : c, DIM == 3 ? c : 0);
  const unsigned rb = row_base<DIM>(g, r, 1, 1);
  // the slot mask is loaded before the pulls and the slot data right after
  // them: a wave with particle nodes has both in flight with the populations
  const uint8_t m = a.mask[rb + (unsigned)c];
  double f[Q];
  gather_row<DIM, Q, ST_F64, false, POST ? 1 : 0, ALN>(a, r, c, f);
  SlotsF64 S;
  load_slots_f64(a, rb + (unsigned)c, m, S);
  if (__any(face)) apply_faces<Q, double>(f, a.faces, a.nfaces, r.i, DIM == 3 ? r.j : c,
                                          DIM == 3 ? c : 0, 0.0);
  collide_f64<DIM, Q, FAX, ALN>(a, f, rb, (unsigned)c, S);
}

template <int DIM, int Q, int FAX, bool ALN>
__device__ __forceinline__ void collide_f64(const CollideArgs& a, double* f, unsigned rb,
                                            unsigned uc, SlotsF64& S) {
  const Geom& g = a.g;
  const int64_t plane = g.plane;
  constexpr bool AL = ALN && !kPlanar<ST_F64>;
  const unsigned oc = AL ? col_off<ST_F64, Q>(uc) : 0u;
  auto dst = [&]<int q>() __attribute__((always_inline)) {
    if constexpr (AL) return pop_ptr_al_w<ST_F64, Q, q>(a.fout, rb, oc);
    else return pop_ptr_w<ST_F64, Q>(a.fout, q, plane, rb, uc);
  };
  const double omega = a.omega, inv2 = a.invCs2, inv4 = a.invCs4;
  // The reference's quadratic term (e.u)(e.u) 0.5 invCs4 (fluid) or
  // 0.5 (e.u)(e.u) invCs4 (solid) halves a product exactly, so it equals
  // (e.u)(e.u) (0.5 invCs4) bit for bit (a difference could only arise
  // below 1e-300, where the term vanishes next to 1.0 anyway).
  const double hinv4 = a.hInvCs4;
  const double F0 = a.F[0], F1 = a.F[1], F2 = a.F[2];
  {
    const unsigned n = rb + uc;
    const uint8_t m = S.m;
    const bool psm = S.psm, o1 = S.o1, o2 = S.o2;
    double B1 = S.B1, B2 = S.B2;
    // scalar copies (kept in registers across the two forms of the particle
    // pass below)
    const double us1[3] = {S.us1[0], S.us1[1], S.us1[2]};
    const double us2[3] = {S.us2[0], S.us2[1], S.us2[2]};
    double Fa[6] = {S.Fa[0], S.Fa[1], S.Fa[2], S.Fa[3], S.Fa[4], S.Fa[5]};
    double rho, jx, jy, jz;
    moments<Q, double>(f, rho, jx, jy, jz);
    jx = jx + F0 * 0.5;
    jy = jy + F1 * 0.5;
    jz = jz + F2 * 0.5;
    const double ux = jx / rho, uy = jy / rho, uz = jz / rho;
    if (a.rho_out) {
      a.rho_out[n] = rho;
      a.u_out[n] = ux;
      a.u_out[plane + n] = uy;
      a.u_out[2 * plane + n] = uz;
    }
    const double usqt = 0.5 * (ux * ux + uy * uy + uz * uz) * inv2;

    // A wave with any particle node runs the particle form for all its
    // lanes: with both slots empty it yields the fluid value (the extra
    // terms add +0.0).  The branch is wave-uniform, so every population
    // store below is issued once by all 64 lanes together.
    if (!psm) {
      Unroll<0, Q>::run([&]<int q>() {
        constexpr double w = VelSet<Q>::w[q];
        const double eu = edot<Q, q>(ux, uy, uz);
        const double feq = rho * w * (1.0 + eu * inv2 + eu * eu * hinv4 - usqt);
        double out = f[q] + (feq - f[q]) * omega;
        if (FAX >= 0 || a.guo)
          out = out + a.guo_pref * guo_f64<Q, q, FAX>(ux, uy, uz, eu, F0, F1, F2, inv2, inv4);
        st<ST_F64>(dst.template operator()<q>(), out, 1.0f);
      });
    } else {
      const double u1sq = 0.5 * (us1[0] * us1[0] + us1[1] * us1[1] + us1[2] * us1[2]) * inv2;
      const double u2sq = 0.5 * (us2[0] * us2[0] + us2[1] * us2[1] + us2[2] * us2[2]) * inv2;
      const double Btot0 = B1 + B2;
      double Btot = Btot0;
      if (Btot0 > 1.0) {
        B1 = B1 / Btot0;
        B2 = B2 / Btot0;
        Btot = 1.0;
      }
      const double oneMinusOmega = 1.0 - omega;
      // TWO: the second slot's terms are formed.  Without LBMPSM_F64_SPLIT
      // every particle wave forms them (zero on lanes without a second
      // particle); with it, a wave in which no lane holds a second particle
      // leaves out the reference's B2 * solid_coll2 = +0.0 (adding it could
      // only change the sign of a zero).
      auto pass = [&]<bool TWO>() __attribute__((always_inline)) {
        Unroll<0, Q>::run([&]<int q>() {
          constexpr double w = VelSet<Q>::w[q];
          constexpr int cx = VelSet<Q>::ex[q], cy = VelSet<Q>::ey[q], cz = VelSet<Q>::ez[q];
          const double fi = f[q];
          const double eu = edot<Q, q>(ux, uy, uz);
          const double feq = rho * w * (1.0 + eu * inv2 + eu * eu * hinv4 - usqt);
          // fi - feq is exactly -(feq - fi) (IEEE subtraction is
          // antisymmetric up to the sign of a zero), so the solid terms
          // reuse the fluid term's difference
          const double dneq = feq - fi;
          // the reference's fi + (1 - Btot) dneq omega + B1 sc1 + B2 sc2,
          // each slot's term added only where the slot is held: an empty
          // slot has B = 0 and sc = 0, whose +0.0 could only change the
          // sign of a zero
          double out = fi + (1.0 - Btot) * dneq * omega;
          // Node force on each slot, src/LBM_PSM_BGK_dynamics.h:101-113.
          if (o1) {
            const double e1 = edot<Q, q>(us1[0], us1[1], us1[2]);
            const double fs = rho * w * (1.0 + e1 * inv2 + e1 * e1 * hinv4 - u1sq);
            const double sc1 = fs - fi + oneMinusOmega * -dneq;
            out = out + B1 * sc1;
            const double t = -B1 * sc1;
            if constexpr (cx > 0) Fa[0] = Fa[0] + t;
            if constexpr (cx < 0) Fa[0] = Fa[0] - t;
            if constexpr (cy > 0) Fa[1] = Fa[1] + t;
            if constexpr (cy < 0) Fa[1] = Fa[1] - t;
            if constexpr (cz > 0) Fa[2] = Fa[2] + t;
            if constexpr (cz < 0) Fa[2] = Fa[2] - t;
          }
          if constexpr (TWO) {
            if (o2) {
              const double e2 = edot<Q, q>(us2[0], us2[1], us2[2]);
              const double fs = rho * w * (1.0 + e2 * inv2 + e2 * e2 * hinv4 - u2sq);
              const double sc2 = fs - fi + oneMinusOmega * -dneq;
              out = out + B2 * sc2;
              const double t = -B2 * sc2;
              if constexpr (cx > 0) Fa[3] = Fa[3] + t;
              if constexpr (cx < 0) Fa[3] = Fa[3] - t;
              if constexpr (cy > 0) Fa[4] = Fa[4] + t;
              if constexpr (cy < 0) Fa[4] = Fa[4] - t;
              if constexpr (cz > 0) Fa[5] = Fa[5] + t;
              if constexpr (cz < 0) Fa[5] = Fa[5] - t;
            }
          }
          if (FAX >= 0 || a.guo)
            out = out + a.guo_pref * guo_f64<Q, q, FAX>(ux, uy, uz, eu, F0, F1, F2, inv2, inv4);
          st<ST_F64>(dst.template operator()<q>(), out, 1.0f);
        });
      };
#if LBMPSM_F64_SPLIT
      if (__any(o2)) pass.template operator()<true>();
      else pass.template operator()<false>();
#else
      pass.template operator()<true>();
#endif
      if (o1) {
        a.Fh[n] = Fa[0];
        a.Fh[plane + n] = Fa[1];
        a.Fh[2 * plane + n] = Fa[2];
      }
      if (o2) {
        a.Fh[3 * plane + n] = Fa[3];
        a.Fh[4 * plane + n] = Fa[4];
        a.Fh[5 * plane + n] = Fa[5];
      }
      if (m & (kMaskFresh0 | kMaskFresh1)) a.mask[n] = m & (kMaskOcc0 | kMaskOcc1);
    }
  }
}

// ---------------------------------------------------------------------------
// FP32 arithmetic on shifted populations g = f - w*rho0 (F32 / F16 storage).
// ---------------------------------------------------------------------------
// Slot data of a node for the FP32 particle form (solid fractions and node
// velocities in float); psm: some lane of the wave holds a particle.
struct SlotsF32 {
  uint8_t m;
  bool psm, o1, o2;
  float B1, B2;
  float us1[3], us2[3];
};

__device__ __forceinline__ void load_slots_f32(const CollideArgs& a, unsigned n, uint8_t m,
                                               SlotsF32& S) {
  const int64_t plane = a.g.plane;
  S.m = m;
  S.psm = __any(m != 0);
  S.o1 = m & 1;
  S.o2 = m & 2;
  S.B1 = S.B2 = 0.0f;
  for (int d = 0; d < 3; ++d) S.us1[d] = S.us2[d] = 0.0f;
  if (!S.psm) return;
  if (S.o1) {
    S.B1 = a.sff[n];
    S.us1[0] = a.usf[n];
    S.us1[1] = a.usf[plane + n];
    S.us1[2] = a.usf[2 * plane + n];
  }
  if (S.o2) {
    S.B2 = a.sff[plane + n];
    S.us2[0] = a.usf[3 * plane + n];
    S.us2[1] = a.usf[4 * plane + n];
    S.us2[2] = a.usf[5 * plane + n];
  }
}

template <int DIM, int Q, int ST, int FAX, bool ALN>
__device__ __forceinline__ void collide_lowp(const CollideArgs& a, float* f, unsigned rb,
                                             unsigned uc, const SlotsF32& S);

// The bracket of the FP32 Guo term, 3 (e - u).F + 9 (e.u)(e.F); FAX >= 0:
// only F[FAX] is nonzero, and the other components' terms (exact zeros) are
// left out.
template <int Q, int q, int FAX>
__device__ __forceinline__ float guo_lowp(float ux, float uy, float uz, float eu, float F0,
                                          float F1, float F2) {
#pragma clang fp contract(fast)
  constexpr int cx = VelSet<Q>::ex[q], cy = VelSet<Q>::ey[q], cz = VelSet<Q>::ez[q];
  constexpr float ex = cx, ey = cy, ez = cz;
  if constexpr (FAX < 0) {
    const float eF = edot<Q, q>(F0, F1, F2);
    return 3.0f * ((ex - ux) * F0 + (ey - uy) * F1 + (ez - uz) * F2) + 9.0f * eu * eF;
  } else {
    constexpr int ca = FAX == 0 ? cx : (FAX == 1 ? cy : cz);
    constexpr float ea = ca;
    const float Fa = FAX == 0 ? F0 : (FAX == 1 ? F1 : F2);
    const float ua = FAX == 0 ? ux : (FAX == 1 ? uy : uz);
    if constexpr (ca != 0) return 3.0f * ((ea - ua) * Fa) + 9.0f * eu * eterm<ca>(Fa);
    else return 3.0f * ((ea - ua) * Fa);
  }
}

// Weights of the FP32 form's even terms (w (drho + rho (...)), the solid
// equilibrium and the Guo term).  The nearest floats to 1/3, 1/18, 1/36 sum
// to 1 + 1.5e-8 (to 4/9, 1/9, 1/36: 1 + 7.4e-9), and a collision then hands
// each node omega * 1.5e-8 * drho of new mass per step: with the pressure
// drop of a packed column (drho ~ 0.1) that is a source spread over the
// whole bed, and the flux it adds at the top moves every drag by ~1e-5 over
// 10^4 steps.  These floats sum to exactly 1: the axis weight is 1/18 (1/9)
// with its mantissa rounded down to a multiple of 4 (6e-8 low), the
// diagonal one half (a quarter) of it, and the rest weight 1 minus the
// others, which is a float.  The odd terms keep their own exact split.
template <int Q>
struct LowpW;
template <>
struct LowpW<19> {
  static constexpr float axis = 14913080.0f / 268435456.0f;  // 2^-28 units
  static constexpr float diag = axis * 0.5f;
  static constexpr float rest = 11184812.0f / 33554432.0f;   // 2^-25 units
  static_assert((double)rest + 6.0 * (double)axis + 12.0 * (double)diag == 1.0);
};
template <>
struct LowpW<9> {
  static constexpr float axis = 14913080.0f / 134217728.0f;  // 2^-27 units
  static constexpr float diag = axis * 0.25f;
  static constexpr float rest = 14913082.0f / 33554432.0f;   // 2^-25 units
  static_assert((double)rest + 4.0 * (double)axis + 4.0 * (double)diag == 1.0);
};
template <int Q, int q>
__device__ __forceinline__ constexpr float lowp_w() {
  constexpr int n2 = VelSet<Q>::ex[q] * VelSet<Q>::ex[q] + VelSet<Q>::ey[q] * VelSet<Q>::ey[q] +
                     VelSet<Q>::ez[q] * VelSet<Q>::ez[q];
  return n2 == 0 ? LowpW<Q>::rest : (n2 == 1 ? LowpW<Q>::axis : LowpW<Q>::diag);
}
// rho times the weight of direction q, from the node's three products.
template <int Q, int q>
__device__ __forceinline__ float lowp_rw(const float (&rw)[3]) {
  constexpr int n2 = VelSet<Q>::ex[q] * VelSet<Q>::ex[q] + VelSet<Q>::ey[q] * VelSet<Q>::ey[q] +
                     VelSet<Q>::ez[q] * VelSet<Q>::ez[q];
  return rw[n2];
}
// Even part of a shifted equilibrium plus the rest r of the population:
// w drho + (rho w t + r), with the product w drho left unrounded inside the
// last fused multiply-add.  Rounded on its own, w drho takes only a few
// phases against the result's grid (the rest weight's mantissa is ~2/3 of
// a power of two, so w drho lands on thirds of a unit), which at drho =
// 0.1 biases it by -0.07 units, a mass loss of 3e-10 per node and step;
// the addend's arbitrary low bits leave the single rounding unbiased.
__device__ __forceinline__ float lowp_even(float w, float drho, float rw, float t, float r) {
  return __builtin_fmaf(w, drho, __builtin_fmaf(rw, t, r));
}

// Odd part of the FP32 equilibrium, 3 w_q (e_q . p) with p = rho u, formed
// as a signed sum of A = 3 w_axis p per axis (w_axis = 1/18 in D3Q19, 1/9 in
// D2Q9); a diagonal weight is w_axis/2 (D3Q19) or w_axis/4 (D2Q9), a power
// of two.  Two roundings would bias the momentum that every step hands to
// the flow, and over thousands of steps under a small body force either one
// moves the drag by ~1e-5:
//  * the float weight alone scales every equilibrium's momentum by one
//    constant factor (1 + 7.4e-9 for w = 1/18), a force omega * 7.4e-9 *
//    rho u per step; 3 w_axis is therefore applied as two floats, hi + lo;
//  * rounding p = j + F/2 to float adds the same error, (F/2 mod ulp(j)), at
//    every node whose j lies in one binade, a fixed fraction of F (up to
//    ~1e-5 of it); the half force is instead folded into the last fused
//    multiply-add, whose single rounding falls on the exact product j * hi.
template <int Q>
struct OddK {
  static constexpr double k = 3.0 * VelSet<Q>::w[1];
  static constexpr float hi = (float)k;
  static constexpr float lo = (float)(k - (double)hi);
};
// A = 3 w_axis (j + F/2), with hf = 3 w_axis F/2 (a per-launch constant).
template <int Q>
__device__ __forceinline__ float odd_axis_j(float j, float hf) {
  return __builtin_fmaf(j, OddK<Q>::hi, __builtin_fmaf(j, OddK<Q>::lo, hf));
}
template <int Q, int q>
__device__ __forceinline__ float eq_odd(float Ax, float Ay, float Az) {
  constexpr int cx = VelSet<Q>::ex[q], cy = VelSet<Q>::ey[q], cz = VelSet<Q>::ez[q];
  constexpr int n2 = cx * cx + cy * cy + cz * cz;
  static_assert(n2 <= 2, "velocity sets with |e|^2 <= 2 only");
  if constexpr (n2 == 0) {
    return 0.0f;
  } else if constexpr (n2 == 1) {
    return edot<Q, q>(Ax, Ay, Az);
  } else {
    constexpr float r = (float)(VelSet<Q>::w[q] / VelSet<Q>::w[1]);
    return r * edot<Q, q>(Ax, Ay, Az);
  }
}

template <int DIM, int Q, bool EXACT, int FAX, typename Emit>
__device__ __forceinline__ void collide_lowp_node(const CollideArgs& a, float* f, unsigned n,
                                                  const SlotsF32& S, Emit&& emit);

template <int DIM, int Q, int ST, bool POST, int FAX, bool ALN>
__global__ __launch_bounds__(256) LBMPSM_LOWP_OCC void k_collide_lowp(const CollideArgs a) {
  const Geom& g = a.g;
  Row r;
  if (!make_row<DIM>(g, bulk_wave(), r)) return;
  const int c = r.c0 + (int)(threadIdx.x & 63);
  if (c >= r.ncols) return;
  const bool face = !POST && a.nfaces &&
      on_face(a.faces, a.nfaces, r.i, DIM == 3 ? r.j : c, DIM == 3 ? c : 0);
  const unsigned rb = row_base<DIM>(g, r, 1, 1);
  // slot mask before the pulls, slot data right after them (as in FP64)
  const uint8_t m = a.mask[rb + (unsigned)c];
  float f[Q];
  gather_row<DIM, Q, ST, false, POST ? 1 : 0, ALN>(a, r, c, f);
  SlotsF32 S;
  load_slots_f32(a, rb + (unsigned)c, m, S);
  if (__any(face)) apply_faces<Q, float>(f, a.faces, a.nfaces, r.i, DIM == 3 ? r.j : c,
                                         DIM == 3 ? c : 0, a.lowp_c[1]);
  collide_lowp<DIM, Q, ST, FAX, ALN>(a, f, rb, (unsigned)c, S);
}

template <int DIM, int Q, int ST, int FAX, bool ALN>
__device__ __forceinline__ void collide_lowp(const CollideArgs& a, float* f, unsigned rb,
                                             unsigned uc, const SlotsF32& S) {
  const int64_t plane = a.g.plane;
  constexpr bool AL = ALN && !kPlanar<ST>;
  const unsigned oc = AL ? col_off<ST, Q>(uc) : 0u;
  collide_lowp_node<DIM, Q, ST == ST_F32, FAX>(a, f, rb + uc, S, [&](auto qc, float out) {
    constexpr int q = decltype(qc)::value;
    if constexpr (AL) st<ST>(pop_ptr_al_w<ST, Q, q>(a.fout, rb, oc), out, a.scale);
    else st<ST>(pop_ptr_w<ST, Q>(a.fout, q, plane, rb, uc), out, a.scale);
  });
}

// One node in FP32: f holds the node's post-stream populations (after the
// faces); emit(q, value) receives each post-collision population in q order
// and the node's force goes to Fh.  EXACT (FP32 storage) forms the even
// terms with w drho unrounded (lowp_even); FP16 storage, whose own rounding
// is 2^13 times coarser, keeps the plain form w (drho + rho t), which needs
// fewer registers.
template <int DIM, int Q, bool EXACT, int FAX, typename Emit>
__device__ __forceinline__ void collide_lowp_node(const CollideArgs& a, float* f, unsigned n,
                                                  const SlotsF32& S, Emit&& emit) {
#pragma clang fp contract(fast)
  const Geom& g = a.g;
  const int64_t plane = g.plane;
  const float omega = a.lowp_c[0];
  const float rho0 = a.lowp_c[1];
  const float F0 = a.lowp_c[2], F1 = a.lowp_c[3], F2 = a.lowp_c[4];
  const float hF0 = a.lowp_c[5], hF1 = a.lowp_c[6], hF2 = a.lowp_c[7];
  const float gp = a.lowp_c[8];
  {
    float drho, jx, jy, jz;
    moments<Q, float>(f, drho, jx, jy, jz);
    const float rho = rho0 + drho;
    const float ir = 1.0f / rho;
    const float ux = (jx + 0.5f * F0) * ir, uy = (jy + 0.5f * F1) * ir,
                uz = (jz + 0.5f * F2) * ir;
    if (a.rho_out) {
      a.rho_out[n] = (double)rho0 + (double)drho;
      a.u_out[n] = ux;
      a.u_out[plane + n] = uy;
      a.u_out[2 * plane + n] = uz;
    }
    const float usq = 1.5f * (ux * ux + uy * uy + uz * uz);
    const float rw[3] = {rho * LowpW<Q>::rest, rho * LowpW<Q>::axis, rho * LowpW<Q>::diag};
    const uint8_t m = S.m;
    // g_eq = w (drho + rho (4.5 (e.u)^2 - 1.5 u^2)) + 3 w e.(rho u), with
    // rho u = j + F/2 (u itself only enters terms even in e); mixed waves
    // run the particle form for all lanes
    // (wave-uniform branch, whole-wave stores)
    const float Ax = odd_axis_j<Q>(jx, hF0), Ay = odd_axis_j<Q>(jy, hF1),
                Az = odd_axis_j<Q>(jz, hF2);
    if (!S.psm) {
      Unroll<0, Q>::run([&]<int q>() {
        constexpr float w = lowp_w<Q, q>();
        const float eu = edot<Q, q>(ux, uy, uz);
        const float geq =
            EXACT ? lowp_even(w, drho, lowp_rw<Q, q>(rw), 4.5f * eu * eu - usq,
                              eq_odd<Q, q>(Ax, Ay, Az))
                  : w * (drho + rho * (4.5f * eu * eu - usq)) + eq_odd<Q, q>(Ax, Ay, Az);
        float out = f[q] + omega * (geq - f[q]);
        if (FAX >= 0 || a.guo) out += gp * w * guo_lowp<Q, q, FAX>(ux, uy, uz, eu, F0, F1, F2);
        emit(std::integral_constant<int, q>{}, out);
      });
    } else {
      const bool o1 = S.o1, o2 = S.o2;
      float B1 = S.B1, B2 = S.B2;
      const float* us1 = S.us1;
      const float* us2 = S.us2;
      float Fa[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
      const float u1sq = 1.5f * (us1[0] * us1[0] + us1[1] * us1[1] + us1[2] * us1[2]);
      const float u2sq = 1.5f * (us2[0] * us2[0] + us2[1] * us2[1] + us2[2] * us2[2]);
      float Btot = B1 + B2;
      if (Btot > 1.0f) {
        B1 = B1 / Btot;
        B2 = B2 / Btot;
        Btot = 1.0f;
      }
      const float omo = 1.0f - omega;
      // TWO: some lane of the wave holds a second particle.  Lanes without
      // one have B2 = 0, so the second-slot terms only add zeros; a wave
      // without any (all but the few waves over near-contacts) leaves them
      // out instead of issuing them under an empty lane mask.
      auto pass = [&]<bool TWO>() {
        Unroll<0, Q>::run([&]<int q>() {
          constexpr float w = lowp_w<Q, q>();
          constexpr int cx = VelSet<Q>::ex[q], cy = VelSet<Q>::ey[q], cz = VelSet<Q>::ez[q];
          const float gi = f[q];
          const float eu = edot<Q, q>(ux, uy, uz);
          const float geq =
              EXACT ? lowp_even(w, drho, lowp_rw<Q, q>(rw), 4.5f * eu * eu - usq,
                                eq_odd<Q, q>(Ax, Ay, Az))
                    : w * (drho + rho * (4.5f * eu * eu - usq)) + eq_odd<Q, q>(Ax, Ay, Az);
          float sc1 = 0.f, sc2 = 0.f;
          // g_s - g_i of the solid equilibrium in one rounding (the addend
          // -g_i keeps it unbiased for a particle at rest too); its momentum
          // takes the even weights, a bias of 6e-8 rho u_s, ~1e-7 of the
          // node's exchange B rho (u_s - u) for particle speeds up to the fluid's
          if (o1) {
            const float e1 = edot<Q, q>(us1[0], us1[1], us1[2]);
            const float ds =
                EXACT ? lowp_even(w, drho, lowp_rw<Q, q>(rw), 3.0f * e1 + 4.5f * e1 * e1 - u1sq,
                                  -gi)
                      : w * (drho + rho * (3.0f * e1 + 4.5f * e1 * e1 - u1sq)) - gi;
            sc1 = ds + omo * (gi - geq);
          }
          if constexpr (TWO) {
            if (o2) {
              const float e2 = edot<Q, q>(us2[0], us2[1], us2[2]);
              const float ds =
                  EXACT ? lowp_even(w, drho, lowp_rw<Q, q>(rw), 3.0f * e2 + 4.5f * e2 * e2 - u2sq,
                                    -gi)
                        : w * (drho + rho * (3.0f * e2 + 4.5f * e2 * e2 - u2sq)) - gi;
              sc2 = ds + omo * (gi - geq);
            }
          }
          float out = gi + (1.0f - Btot) * omega * (geq - gi) + B1 * sc1;
          if constexpr (TWO) out = out + B2 * sc2;
          if (FAX >= 0 || a.guo) out += gp * w * guo_lowp<Q, q, FAX>(ux, uy, uz, eu, F0, F1, F2);
          emit(std::integral_constant<int, q>{}, out);
          const float t1 = -B1 * sc1;
          if constexpr (cx != 0) Fa[0] += cx * t1;
          if constexpr (cy != 0) Fa[1] += cy * t1;
          if constexpr (cz != 0) Fa[2] += cz * t1;
          if constexpr (TWO) {
            const float t2 = -B2 * sc2;
            if constexpr (cx != 0) Fa[3] += cx * t2;
            if constexpr (cy != 0) Fa[4] += cy * t2;
            if constexpr (cz != 0) Fa[5] += cz * t2;
          }
        });
      };
      if (__any(o2)) pass.template operator()<true>();
      else pass.template operator()<false>();
      // node forces continue from the stored value, or from zero for a slot
      // the map set this step
      if (o1) {
        const bool fr = m & kMaskFresh0;
        a.Fhf[n] = (fr ? 0.0f : a.Fhf[n]) + Fa[0];
        a.Fhf[plane + n] = (fr ? 0.0f : a.Fhf[plane + n]) + Fa[1];
        a.Fhf[2 * plane + n] = (fr ? 0.0f : a.Fhf[2 * plane + n]) + Fa[2];
      }
      if (o2) {
        const bool fr = m & kMaskFresh1;
        a.Fhf[3 * plane + n] = (fr ? 0.0f : a.Fhf[3 * plane + n]) + Fa[3];
        a.Fhf[4 * plane + n] = (fr ? 0.0f : a.Fhf[4 * plane + n]) + Fa[4];
        a.Fhf[5 * plane + n] = (fr ? 0.0f : a.Fhf[5 * plane + n]) + Fa[5];
      }
      if (m & (kMaskFresh0 | kMaskFresh1)) a.mask[n] = m & (kMaskOcc0 | kMaskOcc1);
    }
  }
}

// Generic per-node pull used by the read-back kernels (not on the hot path).
template <int DIM, int Q, int ST>
__device__ __forceinline__ void gather(const CollideArgs& a, int i, int j, int k,
                                       typename StoreT<ST>::C* f) {
  Row r;
  const int64_t row = DIM == 3 ? (int64_t)i * a.g.sy + j : (int64_t)i;
  const int ncols = DIM == 3 ? a.g.sz : a.g.sy;
  const int c = DIM == 3 ? k : j;
  make_row<DIM>(a.g, row * ((ncols + 63) / 64) + c / 64, r);
  gather_row<DIM, Q, ST, true>(a, r, c, f);
}

// ---------------------------------------------------------------------------
// Read-back helpers.
// ---------------------------------------------------------------------------
// Global interior coordinate of a stored node, or false for halo nodes.
__device__ __forceinline__ bool interior(const Geom& g, int i, int j, int k, int& gi, int& gj,
                                         int& gk) {
  gi = i + g.off[0] - 1;
  gj = j + g.off[1] - 1;
  gk = (g.dim == 3) ? k + g.off[2] - 1 : 0;
  return gi >= 0 && gi < g.nx && gj >= 0 && gj < g.ny && gk >= 0 && gk < g.nz;
}

template <int DIM, int Q, int ST>
__global__ __launch_bounds__(256) void k_finish(const CollideArgs a, double* out) {
  const Geom& g = a.g;
  const int64_t n = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (n >= g.nodes) return;
  const int k = (int)(n % g.sz);
  const int j = (int)((n / g.sz) % g.sy);
  const int i = (int)(n / ((int64_t)g.sz * g.sy));
  int gi, gj, gk;
  if (!interior(g, i, j, k, gi, gj, gk)) return;
  typename StoreT<ST>::C f[Q];
  gather<DIM, Q, ST>(a, i, j, k, f);
  const int64_t o = (((int64_t)gi * g.ny + gj) * g.nz + gk) * Q;
  Unroll<0, Q>::run([&]<int q>() {
    if constexpr (ST == ST_F64) out[o + q] = f[q];
    else out[o + q] = (double)f[q] + VelSet<Q>::w[q] * a.rho0;
  });
}

// The reference's own lattice layout (src/LBM_PSM_lattice.h:88-90): one halo
// node on both sides of every axis of the active dimension, extents nx+2,
// ny+2 and nz+2 (1 in 2D), populations fastest.  A non-periodic axis stores
// its halo, so its reference coordinate is the stored one; on a periodic
// axis the halo positions show the periodic image, as the reference's halo
// refresh fills them.
__device__ __forceinline__ int ref_extent_axis(const Geom& g, int axis) {
  if (axis == 2 && g.dim == 2) return 1;
  return (axis == 0 ? g.nx : (axis == 1 ? g.ny : g.nz)) + 2;
}
__device__ __forceinline__ int ref_to_stored(const Geom& g, int axis, int R, bool& image) {
  image = false;
  if ((axis == 2 && g.dim == 2) || !g.per[axis]) return R;
  const int n = axis == 0 ? g.nx : (axis == 1 ? g.ny : g.nz);
  image = R == 0 || R == n + 1;
  const int s = R - 1;
  return s < 0 ? s + n : (s >= n ? s - n : s);
}

template <int DIM, int Q, int ST>
__global__ __launch_bounds__(256) void k_finish_ref(const CollideArgs a, double* out) {
  const Geom& g = a.g;
  const int ex = ref_extent_axis(g, 0), ey = ref_extent_axis(g, 1), ez = ref_extent_axis(g, 2);
  const int64_t R = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (R >= (int64_t)ex * ey * ez) return;
  const int K = (int)(R % ez), J = (int)((R / ez) % ey), I = (int)(R / ((int64_t)ey * ez));
  bool m0, m1, m2;
  const int i = ref_to_stored(g, 0, I, m0), j = ref_to_stored(g, 1, J, m1);
  const int k = ref_to_stored(g, 2, K, m2);
  typename StoreT<ST>::C f[Q];
  gather<DIM, Q, ST>(a, i, j, k, f);
  Unroll<0, Q>::run([&]<int q>() {
    if constexpr (ST == ST_F64) out[R * Q + q] = f[q];
    else out[R * Q + q] = (double)f[q] + VelSet<Q>::w[q] * a.rho0;
  });
}

template <int Q, int ST>
__global__ __launch_bounds__(256) void k_set_pop_ref(typename StoreT<ST>::T* f, Geom g,
                                                     const double* in, double rho0, float scale) {
  const int ex = ref_extent_axis(g, 0), ey = ref_extent_axis(g, 1), ez = ref_extent_axis(g, 2);
  const int64_t R = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (R >= (int64_t)ex * ey * ez) return;
  const int K = (int)(R % ez), J = (int)((R / ez) % ey), I = (int)(R / ((int64_t)ey * ez));
  bool m0, m1, m2;
  const int i = ref_to_stored(g, 0, I, m0), j = ref_to_stored(g, 1, J, m1);
  const int k = ref_to_stored(g, 2, K, m2);
  if (m0 || m1 || m2) return;  // periodic halo positions: images of stored nodes
  const int64_t n = ((int64_t)i * g.sy + j) * g.sz + k;
  Unroll<0, Q>::run([&]<int q>() {
    const double v = in[R * Q + q];
    const size_t e = pop_index<ST, Q>(q, g.plane, (unsigned)n);
    if constexpr (ST == ST_F64) f[e] = v;
    else if constexpr (ST == ST_F32) f[e] = (float)(v - VelSet<Q>::w[q] * rho0);
    else f[e] = __float2half_rn((float)(v - VelSet<Q>::w[q] * rho0) * scale);
  });
}

template <int DIM, int Q, int ST>
__global__ __launch_bounds__(256) void k_mass(const CollideArgs a, double* partials) {
  const Geom& g = a.g;
  __shared__ double red[256];
  double acc = 0.0;
  for (int64_t n = (int64_t)blockIdx.x * 256 + threadIdx.x; n < g.nodes;
       n += (int64_t)gridDim.x * 256) {
    const int k = (int)(n % g.sz);
    const int j = (int)((n / g.sz) % g.sy);
    const int i = (int)(n / ((int64_t)g.sz * g.sy));
    int gi, gj, gk;
    if (!interior(g, i, j, k, gi, gj, gk)) continue;
    typename StoreT<ST>::C f[Q];
    gather<DIM, Q, ST>(a, i, j, k, f);
    double s = 0.0;
    Unroll<0, Q>::run([&]<int q>() { s += (double)f[q]; });
    if constexpr (ST != ST_F64) s += a.rho0;
    acc += s;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}

// Fill direction q of every stored node with one value given in storage form.
template <int Q, int ST>
__global__ __launch_bounds__(256) void k_fill(typename StoreT<ST>::T* f, int64_t plane, int q,
                                              double v, float scale) {
  const int64_t n = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (n >= plane) return;
  const size_t e = pop_index<ST, Q>(q, plane, (unsigned)n);
  if constexpr (ST == ST_F64) f[e] = v;
  else if constexpr (ST == ST_F32) f[e] = (float)v;
  else f[e] = __float2half_rn((float)v * scale);
}

template <int Q, int ST>
__global__ __launch_bounds__(256) void k_set_pop(typename StoreT<ST>::T* f, Geom g,
                                                 const double* in, double rho0, float scale) {
  const int64_t n = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (n >= g.nodes) return;
  const int k = (int)(n % g.sz);
  const int j = (int)((n / g.sz) % g.sy);
  const int i = (int)(n / ((int64_t)g.sz * g.sy));
  int gi, gj, gk;
  if (!interior(g, i, j, k, gi, gj, gk)) return;
  const int64_t o = (((int64_t)gi * g.ny + gj) * g.nz + gk) * Q;
  Unroll<0, Q>::run([&]<int q>() {
    const double v = in[o + q];
    const size_t e = pop_index<ST, Q>(q, g.plane, (unsigned)n);
    if constexpr (ST == ST_F64) f[e] = v;
    else if constexpr (ST == ST_F32) f[e] = (float)(v - VelSet<Q>::w[q] * rho0);
    else f[e] = __float2half_rn((float)(v - VelSet<Q>::w[q] * rho0) * scale);
  });
}

// ---------------------------------------------------------------------------
// Launchers.
// ---------------------------------------------------------------------------
static inline int64_t rows_of(const Geom& g) {
  return g.dim == 3 ? (int64_t)g.sx * g.sy : (int64_t)g.sx;
}

// One bulk kernel of the storage form ST (FP64 or reduced).
template <int DIM, int Q, int ST, bool POST, int FAX, bool ALN>
static void launch_bulk(dim3 grid, dim3 block, hipStream_t s, const CollideArgs& a) {
  if constexpr (ST == ST_F64)
    hipLaunchKernelGGL((k_collide_f64<DIM, Q, POST, FAX, ALN>), grid, block, 0, s, a);
  else
    hipLaunchKernelGGL((k_collide_lowp<DIM, Q, ST, POST, FAX, ALN>), grid, block, 0, s, a);
}

template <int DIM, int Q, int ST, bool ALN>
static void launch_fax(int fax, dim3 grid, dim3 block, hipStream_t s, const CollideArgs& a) {
  if (fax == 0) launch_bulk<DIM, Q, ST, false, 0, ALN>(grid, block, s, a);
  else if (fax == 1) launch_bulk<DIM, Q, ST, false, 1, ALN>(grid, block, s, a);
  else if (DIM == 3 && fax == 2)
    launch_bulk<DIM, Q, ST, false, DIM == 3 ? 2 : -1, ALN>(grid, block, s, a);
  else launch_bulk<DIM, Q, ST, false, -1, ALN>(grid, block, s, a);
}

template <int DIM, int Q, int ST>
static void launch_storage(int fax, bool aln, dim3 grid, dim3 block, hipStream_t s,
                           const CollideArgs& a) {
  if (a.poststream) launch_bulk<DIM, Q, ST, true, -1, false>(grid, block, s, a);
  else if (aln) launch_fax<DIM, Q, ST, true>(fax, grid, block, s, a);
  else launch_fax<DIM, Q, ST, false>(fax, grid, block, s, a);
}

template <int DIM, int Q>
static hipError_t collide_dim(int storage, const CollideArgs& a, hipStream_t s) {
  const int64_t rows = rows_of(a.g);
  const int ncols = DIM == 3 ? a.g.sz : a.g.sy;
  const int64_t waves = rows * ((ncols + 63) / 64);
  // measured: two-wave workgroups help the reduced forms on a large lattice
  // (where the particle kernels beside the collide take a larger share) and
  // cost them a few percent on a small one
  const int rw = storage != ST_F64 && waves >= LBMPSM_LARGE_WAVES ? LBMPSM_ROW_WAVES_LOWP
                                                                  : LBMPSM_ROW_WAVES;
  const int wpb = rw > 0 ? rw : std::min((ncols + 63) / 64, 4);
  dim3 grid((unsigned)((waves + wpb - 1) / wpb)), block((unsigned)(64 * wpb));
  // a body force along one lattice axis takes the kernels formed for it
  int fax = -1;
  if (a.guo) {
    const int nz = (a.F[0] != 0.0) + (a.F[1] != 0.0) + (a.F[2] != 0.0);
    if (nz == 1) fax = a.F[0] != 0.0 ? 0 : (a.F[1] != 0.0 ? 1 : 2);
    if (DIM == 2 && fax == 2) fax = -1;
  }
  // rows a whole number of 64-node blocks long take the split addressing
  // (its lane offsets, below ncols * Q * 8 bytes, must fit 32 bits)
  const bool aln = ncols % 64 == 0 && (int64_t)ncols * Q * 8 < ((int64_t)1 << 32);
  switch (storage) {
    case ST_F64: launch_storage<DIM, Q, ST_F64>(fax, aln, grid, block, s, a); break;
    case ST_F32: launch_storage<DIM, Q, ST_F32>(fax, aln, grid, block, s, a); break;
    case ST_F16: launch_storage<DIM, Q, ST_F16>(fax, aln, grid, block, s, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_collide(int storage, const CollideArgs& a, hipStream_t s) {
  return a.g.dim == 3 ? collide_dim<3, 19>(storage, a, s) : collide_dim<2, 9>(storage, a, s);
}

template <int DIM, int Q>
static hipError_t finish_dim(int storage, const CollideArgs& a, double* out, hipStream_t s) {
  dim3 grid((unsigned)((a.g.nodes + 255) / 256));
  switch (storage) {
    case ST_F64: hipLaunchKernelGGL((k_finish<DIM, Q, ST_F64>), grid, dim3(256), 0, s, a, out); break;
    case ST_F32: hipLaunchKernelGGL((k_finish<DIM, Q, ST_F32>), grid, dim3(256), 0, s, a, out); break;
    case ST_F16: hipLaunchKernelGGL((k_finish<DIM, Q, ST_F16>), grid, dim3(256), 0, s, a, out); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

static inline int64_t ref_nodes(const Geom& g) {
  return (int64_t)(g.nx + 2) * (g.ny + 2) * (g.dim == 3 ? g.nz + 2 : 1);
}

template <int DIM, int Q>
static hipError_t finish_ref_dim(int storage, const CollideArgs& a, double* out, hipStream_t s) {
  dim3 grid((unsigned)((ref_nodes(a.g) + 255) / 256));
  switch (storage) {
    case ST_F64: hipLaunchKernelGGL((k_finish_ref<DIM, Q, ST_F64>), grid, dim3(256), 0, s, a, out); break;
    case ST_F32: hipLaunchKernelGGL((k_finish_ref<DIM, Q, ST_F32>), grid, dim3(256), 0, s, a, out); break;
    case ST_F16: hipLaunchKernelGGL((k_finish_ref<DIM, Q, ST_F16>), grid, dim3(256), 0, s, a, out); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_finish_ref(int storage, const CollideArgs& a, double* out, hipStream_t s) {
  return a.g.dim == 3 ? finish_ref_dim<3, 19>(storage, a, out, s)
                      : finish_ref_dim<2, 9>(storage, a, out, s);
}

template <int Q>
static hipError_t set_ref_q(int storage, void* f, const Geom& g, const double* in, double rho0,
                            float scale, hipStream_t s) {
  dim3 grid((unsigned)((ref_nodes(g) + 255) / 256));
  switch (storage) {
    case ST_F64: hipLaunchKernelGGL((k_set_pop_ref<Q, ST_F64>), grid, dim3(256), 0, s, (double*)f, g, in, rho0, scale); break;
    case ST_F32: hipLaunchKernelGGL((k_set_pop_ref<Q, ST_F32>), grid, dim3(256), 0, s, (float*)f, g, in, rho0, scale); break;
    case ST_F16: hipLaunchKernelGGL((k_set_pop_ref<Q, ST_F16>), grid, dim3(256), 0, s, (__half*)f, g, in, rho0, scale); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_set_populations_ref(int storage, void* f, const Geom& g, const double* in,
                                      double rho0, float scale, hipStream_t s) {
  return g.dim == 3 ? set_ref_q<19>(storage, f, g, in, rho0, scale, s)
                    : set_ref_q<9>(storage, f, g, in, rho0, scale, s);
}

hipError_t launch_finish(int storage, const CollideArgs& a, double* out, hipStream_t s) {
  return a.g.dim == 3 ? finish_dim<3, 19>(storage, a, out, s) : finish_dim<2, 9>(storage, a, out, s);
}

template <int DIM, int Q>
static hipError_t mass_dim(int storage, const CollideArgs& a, double* partials, int nb,
                           hipStream_t s) {
  switch (storage) {
    case ST_F64: hipLaunchKernelGGL((k_mass<DIM, Q, ST_F64>), dim3(nb), dim3(256), 0, s, a, partials); break;
    case ST_F32: hipLaunchKernelGGL((k_mass<DIM, Q, ST_F32>), dim3(nb), dim3(256), 0, s, a, partials); break;
    case ST_F16: hipLaunchKernelGGL((k_mass<DIM, Q, ST_F16>), dim3(nb), dim3(256), 0, s, a, partials); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_mass(int storage, const CollideArgs& a, double* partials, int nb,
                       hipStream_t s) {
  return a.g.dim == 3 ? mass_dim<3, 19>(storage, a, partials, nb, s)
                      : mass_dim<2, 9>(storage, a, partials, nb, s);
}

// cb holds one value per direction in storage form (raw for F64, shifted
// for F32/F16).
template <int Q>
static hipError_t fill_planes_q(int storage, void* f, const Geom& g, const double* cb,
                                float scale, hipStream_t s) {
  dim3 grid((unsigned)((g.plane + 255) / 256));
  for (int q = 0; q < Q; ++q) {
    switch (storage) {
      case ST_F64: hipLaunchKernelGGL((k_fill<Q, ST_F64>), grid, dim3(256), 0, s, (double*)f, g.plane, q, cb[q], scale); break;
      case ST_F32: hipLaunchKernelGGL((k_fill<Q, ST_F32>), grid, dim3(256), 0, s, (float*)f, g.plane, q, cb[q], scale); break;
      case ST_F16: hipLaunchKernelGGL((k_fill<Q, ST_F16>), grid, dim3(256), 0, s, (__half*)f, g.plane, q, cb[q], scale); break;
      default: return hipErrorInvalidValue;
    }
  }
  return hipGetLastError();
}

hipError_t launch_fill_planes(int storage, void* f, const Geom& g, const double* cb,
                              double rho0, float scale, hipStream_t s) {
  (void)rho0;
  return g.q == 19 ? fill_planes_q<19>(storage, f, g, cb, scale, s)
                   : fill_planes_q<9>(storage, f, g, cb, scale, s);
}

template <int Q>
static hipError_t setpop_q(int storage, void* f, const Geom& g, const double* in, double rho0,
                           float scale, hipStream_t s) {
  dim3 grid((unsigned)((g.nodes + 255) / 256));
  switch (storage) {
    case ST_F64: hipLaunchKernelGGL((k_set_pop<Q, ST_F64>), grid, dim3(256), 0, s, (double*)f, g, in, rho0, scale); break;
    case ST_F32: hipLaunchKernelGGL((k_set_pop<Q, ST_F32>), grid, dim3(256), 0, s, (float*)f, g, in, rho0, scale); break;
    case ST_F16: hipLaunchKernelGGL((k_set_pop<Q, ST_F16>), grid, dim3(256), 0, s, (__half*)f, g, in, rho0, scale); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_set_populations(int storage, void* f, const Geom& g, const double* in,
                                  double rho0, float scale, hipStream_t s) {
  return g.q == 19 ? setpop_q<19>(storage, f, g, in, rho0, scale, s)
                   : setpop_q<9>(storage, f, g, in, rho0, scale, s);
}

}  // namespace lbmpsm
