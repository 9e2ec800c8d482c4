// C ABI of the lbmpsm library: context management, host-side set-up that
// mirrors `fix lbm-psm` / `fix lbm-psm-bc`, and the launch sequence of one
// LBM step.  See include/lbmpsm.h for the contract of each entry point.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "lbmpsm.h"
#include "lbmpsm_internal.h"

using namespace lbmpsm;

namespace {

std::string g_create_error;

struct ProfEntry {
  std::string name;
  hipEvent_t a, b;
};

}  // namespace

struct lbmpsm_ctx {
  lbmpsm_config cfg{};
  Geom g{};
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  // a particle list's upload, geometry and solid fractions run on stream2,
  // beside the previous step's collide and reduction on `stream`; the map
  // waits for them (ev_sf), and the next list's preparation waits for the map
  // (ev_mapdone), whose brick lists and codes it reuses
  hipStream_t stream2 = nullptr;
  hipEvent_t ev_sf = nullptr;
  hipEvent_t ev_mapdone = nullptr;
  hipEvent_t ev_collide = nullptr;  // end of the last collide (LBMPSM_SF_PLACEMENT 2)

  // populations: two buffers, cur holds the current state
  void* f[2] = {nullptr, nullptr};
  size_t elem = 8;
  int cur = 0;
  int poststream = 1;
  double cb_raw[19] = {0};     // initial equilibrium (never-written halo values)
  double cb_store[19] = {0};   // same in storage form
  double rho0 = 1.0;
  float scale = 1.0f, inv_scale = 1.0f;

  SlotsDev slots;
  double* rho = nullptr;  // macros [plane], [3][plane]
  double* u = nullptr;
  int have_macros = 0;

  ParticleDev part;
  BrickDev br;
  unsigned short* sfcode = nullptr;  // per-entry zone solid-fraction codes
  int zone_cap = 0;                  // codes per entry
  int zone_span = 0;                 // zone extent bound per axis (nodes)
  size_t sfcode_elems = 0;
  // pinned host staging for one particle upload and the hydro download
  void* h_stage = nullptr;
  size_t h_stage_bytes = 0;
  hipEvent_t stage_done = nullptr;
  double* h_out = nullptr;
  size_t h_out_bytes = 0;
  int hydro_ready = 0;
  int hydro_n = 0;

  // boundary faces
  int bc_keyword = 0;
  int procgrid[3] = {1, 1, 1};  // reference rank grid whose face ranges are emulated
  int nfaces = 0;
  Face faces[kMaxFaces];

  double* d_cb = nullptr;       // [19] cb_store on the device
  Face* d_faces = nullptr;      // [kMaxFaces] on the device
  double* d_scratch = nullptr;  // finish / mass scratch
  size_t d_scratch_bytes = 0;

  int64_t device_bytes = 0;

  // adjoint of a run with frozen particles (lbmpsm_adjoint_*)
  struct Adj {
    int ready = 0;
    int steps = 0, every = 0, ncp = 0, n = 0;
    std::vector<int> post;          // per checkpoint: 1 = post-stream state
    double* ckpt = nullptr;         // [ncp][q*plane] state at each segment start
    double* seg = nullptr;          // [every-1][q*plane] states inside a segment
    double* lam = nullptr;          // [2][q*plane]
    double* acc = nullptr;          // [10][plane]
    double* seed = nullptr;         // [6][plane]
    int* writer = nullptr;          // [2][plane]
    double* bar = nullptr;          // [n][12]
    double* eout = nullptr;         // [n][6]
    double* partials = nullptr;     // [1024][4]
    double* Fh_save = nullptr;      // [6][plane] slot forces of the forward's last step
    uint8_t* mask_save = nullptr;   // [plane] slot masks of the forward's last step
    size_t ckpt_n = 0, seg_n = 0, ent_n = 0;
  } adj;

  // profiling
  int prof = 0;
  std::vector<ProfEntry> pending;
  std::vector<hipEvent_t> free_events;
  std::map<std::string, std::pair<double, int64_t>> totals;
};

namespace {

// Where a list's geometry and solid fractions run: 0 on the main stream
// before its map; 1 on the second stream as soon as the list is given
// (beside the previous step's collide and reduction); 2 on the second stream
// once the previous step's collide has finished (beside its reduction).
#ifndef LBMPSM_SF_PLACEMENT
#define LBMPSM_SF_PLACEMENT 1
#endif
constexpr int kSfPlacement = LBMPSM_SF_PLACEMENT;


int fail(lbmpsm_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  else g_create_error = msg;
  return code;
}

#define HIPCHK(ctx, expr)                                                              \
  do {                                                                                 \
    hipError_t _e = (expr);                                                            \
    if (_e != hipSuccess)                                                              \
      return fail((ctx), LBMPSM_ERR_HIP,                                               \
                  std::string(#expr) + ": " + hipGetErrorString(_e));                  \
  } while (0)

hipEvent_t take_event(lbmpsm_ctx* c) {
  if (!c->free_events.empty()) {
    hipEvent_t e = c->free_events.back();
    c->free_events.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// Brackets a launch with events on the context stream when profiling.
struct ProfScope {
  lbmpsm_ctx* c;
  const char* name;
  hipStream_t st;
  hipEvent_t a = nullptr, b = nullptr;
  bool on = false;
  ProfScope(lbmpsm_ctx* c_, const char* n, hipStream_t s = nullptr)
      : c(c_), name(n), st(s ? s : c_->stream) {
    // mode 1: every kernel; mode 2: the bulk collide only
    on = c->prof == 1 || (c->prof == 2 && std::strcmp(n, "collide_stream") == 0);
    if (on) {
      a = take_event(c);
      b = take_event(c);
      if (a && b) (void)hipEventRecord(a, st);
    }
  }
  ~ProfScope() {
    if (on && a && b) {
      (void)hipEventRecord(b, st);
      c->pending.push_back({name, a, b});
    }
  }
};

// Both streams idle (before buffers they may use are freed or reused).
int sync_all(lbmpsm_ctx* c) {
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (c->stream2) HIPCHK(c, hipStreamSynchronize(c->stream2));
  return LBMPSM_OK;
}

int prof_collect(lbmpsm_ctx* c) {
  if (c->pending.empty()) return LBMPSM_OK;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (auto& p : c->pending) {
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, p.a, p.b));
    auto& t = c->totals[p.name];
    t.first += ms;
    t.second += 1;
    c->free_events.push_back(p.a);
    c->free_events.push_back(p.b);
  }
  c->pending.clear();
  return LBMPSM_OK;
}

template <typename T>
int dev_alloc(lbmpsm_ctx* c, T** p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  void* q = nullptr;
  hipError_t e = hipMalloc(&q, count * sizeof(T));
  if (e != hipSuccess)
    return fail(c, LBMPSM_ERR_NOMEM,
                "device allocation of " + std::to_string(count * sizeof(T)) + " bytes failed: " +
                    hipGetErrorString(e));
  *p = (T*)q;
  c->device_bytes += (int64_t)(count * sizeof(T));
  return LBMPSM_OK;
}

void dev_free(void* p) {
  if (p) (void)hipFree(p);
}

// Reference feq with a node's rho and u, src/LBM_PSM_dynamics.h:43-59.
double feq_ref(int q, int iq, double rho, double ux, double uy, double uz) {
  static const int ex19[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0};
  static const int ey19[19] = {0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1};
  static const int ez19[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, -1, 1, -1, 1};
  static const int ex9[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
  static const int ey9[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
  const double csPow2 = 1.0 / 3.0;
  const double csPow4 = csPow2 * csPow2;
  const double invCsPow2 = 1.0 / csPow2;
  const double invCsPow4 = 1.0 / csPow4;
  double ex_i, ey_i, ez_i, w_i;
  if (q == 19) {
    ex_i = ex19[iq];
    ey_i = ey19[iq];
    ez_i = ez19[iq];
    w_i = iq == 0 ? 1.0 / 3.0 : (iq < 7 ? 1.0 / 18.0 : 1.0 / 36.0);
  } else {
    ex_i = ex9[iq];
    ey_i = ey9[iq];
    ez_i = 0;
    w_i = iq == 0 ? 4.0 / 9.0 : (iq < 5 ? 1.0 / 9.0 : 1.0 / 36.0);
  }
  double eiq_dot_u = ex_i * ux + ey_i * uy + ez_i * uz;
  return rho * w_i *
         (1.0 + eiq_dot_u * invCsPow2 + eiq_dot_u * eiq_dot_u * 0.5 * invCsPow4 -
          0.5 * (ux * ux + uy * uy + uz * uz) * invCsPow2);
}

double wq(int q, int iq) {
  if (q == 19) return iq == 0 ? 1.0 / 3.0 : (iq < 7 ? 1.0 / 18.0 : 1.0 / 36.0);
  return iq == 0 ? 4.0 / 9.0 : (iq < 5 ? 1.0 / 9.0 : 1.0 / 36.0);
}

CollideArgs make_args(lbmpsm_ctx* c, int store_macros) {
  CollideArgs a{};
  a.g = c->g;
  a.fin = c->f[c->cur];
  a.fout = c->f[1 - c->cur];
  a.mask = c->slots.mask;
  a.sf = c->slots.sf;
  a.us = c->slots.us;
  a.Fh = c->slots.Fh;
  a.sff = c->slots.sff;
  a.usf = c->slots.usf;
  a.Fhf = c->slots.Fhf;
  a.rho_out = store_macros ? c->rho : nullptr;
  a.u_out = store_macros ? c->u : nullptr;
  a.poststream = c->poststream;
  const double* F = c->cfg.F_lbm;
  a.guo = (F[0] * F[0] + F[1] * F[1] + F[2] * F[2]) > 0.0 ? 1 : 0;
  a.omega = 1.0 / c->cfg.tau;
  a.guo_pref = 1.0 - 0.5 * a.omega;
  a.F[0] = F[0];
  a.F[1] = F[1];
  a.F[2] = F[2];
  const double csPow2 = 1.0 / 3.0;
  const double csPow4 = csPow2 * csPow2;
  a.invCs2 = 1.0 / csPow2;
  a.invCs4 = 1.0 / csPow4;
  a.hInvCs4 = 0.5 * a.invCs4;
  a.rho0 = c->rho0;
  a.scale = c->scale;
  a.inv_scale = c->inv_scale;
  a.cb = c->d_cb;
  a.cbf = (const float*)(c->d_cb + 19);
  {
    const double k = 3.0 * (c->g.q == 19 ? VelSet<19>::w[1] : VelSet<9>::w[1]);
    const double v[9] = {a.omega, a.rho0, F[0], F[1], F[2], 0.5 * k * F[0], 0.5 * k * F[1],
                         0.5 * k * F[2], a.guo_pref};
    for (int i = 0; i < 9; ++i) a.lowp_c[i] = (float)v[i];
  }
  a.faces = c->d_faces;
  a.nfaces = c->nfaces;
  return a;
}

int ensure_scratch(lbmpsm_ctx* c, size_t bytes) {
  if (c->d_scratch_bytes >= bytes) return LBMPSM_OK;
  if (c->d_scratch) {
    HIPCHK(c, hipStreamSynchronize(c->stream));
    dev_free(c->d_scratch);
    c->device_bytes -= (int64_t)c->d_scratch_bytes;
    c->d_scratch = nullptr;
    c->d_scratch_bytes = 0;
  }
  int rc = dev_alloc(c, (char**)&c->d_scratch, bytes);
  if (rc) return rc;
  c->d_scratch_bytes = bytes;
  return LBMPSM_OK;
}

int ensure_host_out(lbmpsm_ctx* c, size_t bytes) {
  if (c->h_out_bytes >= bytes) return LBMPSM_OK;
  if (c->h_out) (void)hipHostFree(c->h_out);
  c->h_out = nullptr;
  c->h_out_bytes = 0;
  HIPCHK(c, hipHostMalloc((void**)&c->h_out, bytes));
  c->h_out_bytes = bytes;
  return LBMPSM_OK;
}

int ensure_particles(lbmpsm_ctx* c, int n) {
  ParticleDev& P = c->part;
  if (n <= P.cap) return LBMPSM_OK;
  // earlier steps may still read these buffers
  int rc = sync_all(c);
  if (rc) return rc;
  int cap = std::max(n, std::max(64, P.cap * 2));
  void* old[] = {P.raw, P.xl, P.kin, P.zone, P.rzone, P.out, P.hpart,
                 P.raw_alt, P.xl_alt, P.kin_alt, P.zone_alt, P.rzone_alt};
  for (void* p : old) dev_free(p);
  rc |= dev_alloc(c, &P.raw, 84 * (size_t)cap);
  rc |= dev_alloc(c, &P.xl, 4 * (size_t)cap);
  rc |= dev_alloc(c, &P.kin, 6 * (size_t)cap);
  rc |= dev_alloc(c, &P.zone, 6 * (size_t)cap);
  rc |= dev_alloc(c, &P.rzone, 6 * (size_t)cap);
  rc |= dev_alloc(c, &P.raw_alt, 84 * (size_t)cap);
  rc |= dev_alloc(c, &P.xl_alt, 4 * (size_t)cap);
  rc |= dev_alloc(c, &P.kin_alt, 6 * (size_t)cap);
  rc |= dev_alloc(c, &P.zone_alt, 6 * (size_t)cap);
  rc |= dev_alloc(c, &P.rzone_alt, 6 * (size_t)cap);
  rc |= dev_alloc(c, &P.out, 12 * (size_t)cap);
  rc |= dev_alloc(c, &P.hpart, (size_t)kHydroGroups * 12 * (size_t)cap);
  if (rc) return LBMPSM_ERR_NOMEM;
  P.cap = cap;
  if (c->h_stage) (void)hipHostFree(c->h_stage);
  c->h_stage = nullptr;
  HIPCHK(c, hipHostMalloc(&c->h_stage, 84 * (size_t)cap));
  c->h_stage_bytes = 84 * (size_t)cap;
  return LBMPSM_OK;
}

// Codes per entry for k_entry_sf: the largest zone box, (2 r/dx + 8) nodes
// per axis bounds every map zone (src/LBM_PSM_exchangeParticleData.cpp:41-48).
int size_codes(lbmpsm_ctx* c, int n, double rmax) {
  int rc = 0;
  const double span = 2.0 * rmax / c->cfg.dx + 8.0;
  int64_t ext[3];
  const int se[3] = {c->g.sx, c->g.sy, c->g.sz};
  for (int d = 0; d < 3; ++d) ext[d] = std::min<int64_t>((int64_t)span, se[d]);
  int64_t cap = ext[0] * ext[1] * ext[2];
  // zones larger than 2^21 nodes are computed inside the map kernel instead
  cap = std::min<int64_t>(cap, (int64_t)1 << 21);
  const size_t need = (size_t)cap * (size_t)n;
  if (need > c->sfcode_elems) {
    rc = sync_all(c);
    if (rc) return rc;
    const size_t alloc = std::max(need, c->sfcode_elems * 2);
    dev_free(c->sfcode);
    c->device_bytes -= (int64_t)(c->sfcode_elems * sizeof(unsigned short));
    c->sfcode = nullptr;
    c->sfcode_elems = 0;
    rc = dev_alloc(c, &c->sfcode, alloc);
    if (rc) return rc;
    c->sfcode_elems = alloc;
  }
  c->zone_cap = (int)cap;
  c->zone_span = (int)std::min<double>(span, 2147483647.0);
  return LBMPSM_OK;
}

// Face table of `fix lbm-psm-bc`, src/fix_LBM_PSM_BC.cpp:223-401, in the
// reference's halo'd coordinates, converted to stored coordinates.
void add_face_one(lbmpsm_ctx* c, int kind, int axis, int plane_ref, int a_lo, int a_hi, int b_lo,
                  int b_hi, double ux, double uy, double uz, double rho);

// One face given by its range on one rank.  With a reference rank grid
// (c->procgrid), every rank that holds the face applies it over the same
// range of its own local indices (src/fix_LBM_PSM_BC.cpp:223-400 index the
// rank's local lattice), so a range that leaves out interior nodes next to
// the ends leaves them out at every rank boundary too: the face becomes one
// segment per rank along each tangential axis.  Ranks split the interior
// nodes of an axis as src/LBM_PSM_lattice.cpp:34-63 does (the last rank takes
// the remainder).
void add_face(lbmpsm_ctx* c, int kind, int axis, int plane_ref, int a_lo, int a_hi, int b_lo,
              int b_hi, double ux, double uy, double uz, double rho) {
  const Geom& g = c->g;
  const int ta = axis == 0 ? 1 : 0;
  const int tb = axis == 2 ? 1 : 2;
  const int nglob[3] = {g.nx, g.ny, g.dim == 3 ? g.nz : 1};
  // shifts of the range from the interior [1, n] on one rank
  const int sa_lo = a_lo - 1, sa_hi = nglob[ta] - a_hi;
  const int sb_lo = b_lo - 1, sb_hi = nglob[tb] - b_hi;
  const bool b_active = !(g.dim == 2 && tb == 2);
  const int pa = c->procgrid[ta], pb = b_active ? c->procgrid[tb] : 1;
  for (int ra = 0; ra < pa; ++ra) {
    const int la = nglob[ta] / pa + (ra == pa - 1 ? nglob[ta] % pa : 0);
    const int sa = 1 + ra * (nglob[ta] / pa);
    for (int rb = 0; rb < pb; ++rb) {
      int lo_b = b_lo, hi_b = b_hi;
      if (b_active) {
        const int lb = nglob[tb] / pb + (rb == pb - 1 ? nglob[tb] % pb : 0);
        const int sb = 1 + rb * (nglob[tb] / pb);
        lo_b = sb + sb_lo;
        hi_b = sb + lb - 1 - sb_hi;
      }
      add_face_one(c, kind, axis, plane_ref, sa + sa_lo, sa + la - 1 - sa_hi, lo_b, hi_b, ux, uy,
                   uz, rho);
    }
  }
}

void add_face_one(lbmpsm_ctx* c, int kind, int axis, int plane_ref, int a_lo, int a_hi, int b_lo,
                  int b_hi, double ux, double uy, double uz, double rho) {
  if (c->nfaces >= kMaxFaces) return;  // rejected by the caller (face_count)
  const Geom& g = c->g;
  const int ta = axis == 0 ? 1 : 0;         // first tangential axis
  const int tb = axis == 2 ? 1 : 2;         // second tangential axis
  Face F{};
  F.kind = kind;
  F.axis = axis;
  F.plane = plane_ref - g.off[axis];
  F.a_lo = a_lo - g.off[ta];
  F.a_hi = a_hi - g.off[ta];
  F.b_lo = b_lo - g.off[tb];
  F.b_hi = b_hi - g.off[tb];
  F.u[0] = ux;
  F.u[1] = uy;
  F.u[2] = uz;
  F.rho = rho;
  c->faces[c->nfaces++] = F;
}

void build_faces(lbmpsm_ctx* c, int typeBC, double U, double rho_outlet) {
  c->nfaces = 0;
  const int e = 1;
  const int nx = c->g.nx + 2, ny = c->g.ny + 2;  // reference halo'd sizes
  const int nz = c->g.dim == 3 ? c->g.nz + 2 : 1;
  if (c->g.dim == 2) {
    // 2D faces have no z range: node k = 0 only.
    auto xn = [&](double u) { add_face(c, FK_2D_VEL_XN, 0, e, e, ny - 1 - e, 0, 0, u, 0, 0, 0); };
    auto xp = [&](double u) { add_face(c, FK_2D_VEL_XP, 0, nx - 1 - e, e, ny - 1 - e, 0, 0, u, 0, 0, 0); };
    auto dxp = [&](double r) { add_face(c, FK_2D_DENS_XP, 0, nx - 1 - e, e, ny - 1 - e, 0, 0, 0, 0, 0, r); };
    auto yn = [&](int i0, int i1, double u) { add_face(c, FK_2D_VEL_YN, 1, 0 + e, i0, i1, 0, 0, u, 0, 0, 0); };
    auto yp = [&](int i0, int i1, double u) { add_face(c, FK_2D_VEL_YP, 1, ny - 1 - e, i0, i1, 0, 0, u, 0, 0, 0); };
    switch (typeBC) {
      case 1: yn(e, nx - 1 - e, -U); yp(e, nx - 1 - e, U); break;
      case 2: xn(U); dxp(rho_outlet); yn(1 + e, nx - 2 - e, U); yp(1 + e, nx - 2 - e, U); break;
      case 3: yn(e, nx - 1 - e, 0.0); yp(e, nx - 1 - e, 0.0); break;
      case 4: xn(U); dxp(rho_outlet); yn(e, nx - 1 - e, 0.0); yp(e, nx - 1 - e, 0.0); break;
      case 5: xn(0.0); xp(0.0); yn(e, nx - 1 - e, 0.0); yp(e, nx - 1 - e, 0.0); break;
      default: break;
    }
    return;
  }
  auto xn = [&](double ux) { add_face(c, FK_3D_VEL_XN, 0, e, e, ny - 1 - e, e, nz - 1 - e, ux, 0, 0, 0); };
  auto xp = [&](double ux) { add_face(c, FK_3D_VEL_XP, 0, nx - 1 - e, e, ny - 1 - e, e, nz - 1 - e, ux, 0, 0, 0); };
  auto dxp = [&](double r) { add_face(c, FK_3D_DENS_XP, 0, nx - 1 - e, e, ny - 1 - e, e, nz - 1 - e, 0, 0, 0, r); };
  auto yn = [&](double ux) { add_face(c, FK_3D_VEL_YN, 1, e, e, nx - 1 - e, e, nz - 1 - e, ux, 0, 0, 0); };
  auto yp = [&](double ux) { add_face(c, FK_3D_VEL_YP, 1, ny - 1 - e, e, nx - 1 - e, e, nz - 1 - e, ux, 0, 0, 0); };
  auto zn = [&](double ux) { add_face(c, FK_3D_VEL_ZN, 2, e, e, nx - 1 - e, e, ny - 1 - e, ux, 0, 0, 0); };
  auto zp = [&](double ux) { add_face(c, FK_3D_VEL_ZP, 2, nz - 1 - e, e, nx - 1 - e, e, ny - 1 - e, ux, 0, 0, 0); };
  switch (typeBC) {
    case 1: yn(-U); yp(U); break;
    case 2: xn(U); dxp(rho_outlet); yn(U); yp(U); zn(U); zp(U); break;
    case 3: yn(0.0); yp(0.0); break;
    case 4: xn(U); dxp(rho_outlet); yn(0.0); yp(0.0); zn(0.0); zp(0.0); break;
    case 5: xn(0.0); xp(0.0); yn(0.0); yp(0.0); zn(0.0); zp(0.0); break;
    default: break;
  }
}

// The list is read from the context's packed block, or (dev non-null) from
// the caller's device arrays, whose tags are then copied into the block.
struct DevList {
  const int* tag;
  const double *x, *v, *w, *r;
};

int run_map(lbmpsm_ctx* c, const DevList* dev = nullptr) {
  ParticleDev& P = c->part;
  BrickDev& B = c->br;
  const int n = P.n;
  PrepArgs pa{};
  pa.g = c->g;
  pa.n = n;
  if (dev) {
    pa.x = dev->x;
    pa.v = dev->v;
    pa.w = dev->w;
    pa.r = dev->r;
    pa.tag_in = dev->tag;
    pa.tag_out = (int*)P.tag(n);
  } else {
    pa.x = P.x(n);
    pa.v = P.v(n);
    pa.w = P.w(n);
    pa.r = P.r(n);
    pa.tag_in = P.tag(n);
    pa.tag_out = nullptr;
  }
  pa.xl = P.xl;
  pa.kin = P.kin;
  pa.zone = P.zone;
  pa.rzone = P.rzone;
  for (int d = 0; d < 3; ++d) pa.origin[d] = c->cfg.origin[d];
  pa.dx = c->cfg.dx;
  pa.Uc = c->cfg.Uc;
  pa.u_lb = c->cfg.u_lb;
  pa.tc = c->cfg.tc;
  pa.dt_d = c->cfg.dt_d;
  pa.count = B.count;
  pa.list = B.list;
  pa.slow_count = B.slow_count;
  pa.bi = B.bi;
  pa.bj = B.bj;
  pa.bk = B.bk;
  pa.nbi = B.nbi;
  pa.nbj = B.nbj;
  pa.nbk = B.nbk;
  // this list's preparation reuses the brick lists, codes and work-list
  // counter of the previous map: it waits for it (the upload, if any, is
  // already queued behind the same wait)
  hipStream_t sp = c->stream2;
  if (kSfPlacement == 0) {  // on the main stream, after the upload
    HIPCHK(c, hipEventRecord(c->ev_sf, c->stream2));
    HIPCHK(c, hipStreamWaitEvent(c->stream, c->ev_sf, 0));
    sp = c->stream;
  } else if (kSfPlacement == 2) {  // after the last collide
    HIPCHK(c, hipStreamWaitEvent(c->stream2, c->ev_collide, 0));
  }
  {
    ProfScope ps(c, "map_bin", sp);  // counts are zero: set at create, reset by k_map
    HIPCHK(c, launch_prep_bin(pa, sp));
  }
  MapArgs ma{};
  ma.g = c->g;
  ma.s = c->slots;
  ma.n = n;
  ma.tag = P.tag(n);
  ma.xl = P.xl;
  ma.kin = P.kin;
  ma.zone = P.zone;
  ma.list = B.list;
  ma.count = B.count;
  ma.count_reset = B.count;
  ma.bi = B.bi;
  ma.bj = B.bj;
  ma.bk = B.bk;
  ma.nbi = B.nbi;
  ma.nbj = B.nbj;
  ma.nbk = B.nbk;
  ma.hsq_half = c->g.dim == 3 ? std::sqrt(3.1) / 2.0 : std::sqrt(2.1) / 2.0;
  ma.sfcode = c->sfcode;
  ma.zone_cap = c->zone_cap;
  ma.zone_span = c->zone_span;
  ma.slow_list = B.slow_list;
  ma.slow_count = B.slow_count;
  {
    ProfScope ps(c, "map_fraction", sp);
    HIPCHK(c, launch_entry_sf(ma, sp));
  }
  if (kSfPlacement != 0) {
    HIPCHK(c, hipEventRecord(c->ev_sf, c->stream2));
    HIPCHK(c, hipStreamWaitEvent(c->stream, c->ev_sf, 0));
  }
  {
    ProfScope ps(c, "map_particles");
    HIPCHK(c, launch_map(ma, B.nbricks, c->stream));
  }
  HIPCHK(c, hipEventRecord(c->ev_mapdone, c->stream));
  return LBMPSM_OK;
}

// A new list goes into the arrays of the list before the current one (whose
// step is complete once the current list's map is): swap the two sets and
// order stream2 after that map.
int next_list(lbmpsm_ctx* c) {
  ParticleDev& P = c->part;
  std::swap(P.raw, P.raw_alt);
  std::swap(P.xl, P.xl_alt);
  std::swap(P.kin, P.kin_alt);
  std::swap(P.zone, P.zone_alt);
  std::swap(P.rzone, P.rzone_alt);
  HIPCHK(c, hipStreamWaitEvent(c->stream2, c->ev_mapdone, 0));
  return LBMPSM_OK;
}

// Static scratch of lbmpsm_step_noalloc: the per-entry group sums of the
// force/torque/stresslet reduction ([n][kHydroGroups][12] doubles) for
// callers that cannot allocate.  Each device holds its own copy at its own
// address; host code uses that address (noalloc_scratch), never the name.
constexpr int64_t kNoAllocScratchDoubles = (int64_t)1 << 21;  // 16 MiB
static __device__ __align__(256) double scratch[kNoAllocScratchDoubles];
constexpr int kMaxDevices = 64;
double* g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_mutex;

// Once per device, with that device current: the device address of its
// copy of `scratch`.
int noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return LBMPSM_ERR_LIMIT;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (g_scratch_ptr[device]) return LBMPSM_OK;
  double* ptr = nullptr;
  if (hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr)
    return LBMPSM_ERR_HIP;
  g_scratch_ptr[device] = ptr;
  return LBMPSM_OK;
}

double* noalloc_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  return g_scratch_ptr[device];
}

// The launches of one step: collide-stream (faces included) and the
// force/torque/stresslet reduction with its group sums in `part` and its
// per-entry results in `out` ([n][12], device memory).
int step_launch(lbmpsm_ctx* c, int store, double* part, double* out, bool own_out) {
  // the allocation-free form records no profiling events
  struct ProfOff {
    lbmpsm_ctx* c;
    int saved;
    ~ProfOff() { c->prof = saved; }
  } prof_off{c, c->prof};
  if (!own_out) c->prof = 0;
  CollideArgs a = make_args(c, store);
  {
    ProfScope ps(c, "collide_stream");
    HIPCHK(c, launch_collide(c->cfg.storage, a, c->stream));
  }
  if (kSfPlacement == 2) HIPCHK(c, hipEventRecord(c->ev_collide, c->stream));
  c->cur = 1 - c->cur;
  c->poststream = 0;
  if (store) c->have_macros = 1;
  HydroArgs h{};
  h.g = c->g;
  h.n = c->part.n;
  h.tag = c->part.tag(c->part.n);
  h.xl = c->part.xl;
  h.rzone = c->part.rzone;
  h.slot_tag = c->slots.tag;
  h.mask = c->slots.mask;
  h.Fh = c->slots.Fh;
  h.Fhf = c->slots.Fhf;
  h.forceFactor = c->cfg.force_factor;
  h.torqueFactor = c->cfg.torque_factor;
  h.part = part;
  h.out = out;
  if (h.n > 0) {
    ProfScope ps(c, "hydro_reduce");
    HIPCHK(c, launch_hydro(h, c->stream));
  }
  // lbmpsm_get_hydro reads the context's own results only
  c->hydro_ready = own_out ? 1 : 0;
  c->hydro_n = c->part.n;
  return LBMPSM_OK;
}

}  // namespace

extern "C" {

const char* lbmpsm_last_error(const lbmpsm_ctx* ctx) {
  return ctx ? ctx->err.c_str() : g_create_error.c_str();
}

int lbmpsm_config_from_fix(lbmpsm_config* cfg, int dim, int Nlc, double lc, double rho,
                           double nu, double Re, double tau, const double Fext[3],
                           const double boxlo[3], const double prd[3], const int periodic[3]) {
  if (!cfg || !Fext || !boxlo || !prd || !periodic) return fail(nullptr, LBMPSM_ERR_ARG, "null argument");
  if (dim != 2 && dim != 3) return fail(nullptr, LBMPSM_ERR_ARG, "dimension must be 2 or 3");
  if (Nlc <= 1 || lc <= 0.0 || rho <= 0.0 || nu <= 0.0 || Re <= 0.0)
    return fail(nullptr, LBMPSM_ERR_ARG, "Illegal fix lbm-psm command");
  if (tau <= 0.5) return fail(nullptr, LBMPSM_ERR_ARG, "Illegal fix lbm-psm command: tau <= 0.5");
  std::memset(cfg, 0, sizeof(*cfg));
  // UnitConversion, src/LBM_PSM_unitConversion.cpp:15-38
  const double Uc_d = 1.0, Lc_d = 1.0;
  const double csPow2 = 1.0 / 3.0;
  const double nu_lb = (tau - 0.5) * csPow2;
  const double dx = lc / (double)(Nlc - 1);
  const double dx_d = Lc_d / (double)(Nlc - 1);
  const double Uc = Re * nu / lc;
  const double u_lb = Re * nu_lb / (double)(Nlc - 1);
  const double dt_d = dx_d * u_lb / Uc_d;
  const double tc = lc / Uc;
  const double forceFactor = rho * std::pow(lc, 4) / std::pow(tc, 2) * std::pow(dx_d, 4) / std::pow(dt_d, 2);
  const double torqueFactor = rho * std::pow(lc, 5) / std::pow(tc, 2) * std::pow(dx_d, 5) / std::pow(dt_d, 2);
  const double volumeForceFactor = rho * lc / std::pow(tc, 2) * dx_d / std::pow(dt_d, 2);
  // src/fix_LBM_PSM.cpp:122-124
  const double Ma = u_lb / (1.0 / std::sqrt(3.0));
  if (Ma > 0.3)
    return fail(nullptr, LBMPSM_ERR_ARG,
                "Mach number is Ma > 0.3. Aborting simulation. Choose parameters which give Ma < 0.3 for stability reasons.");
  // src/fix_LBM_PSM.cpp:135-140
  const double SMALL = 1e-15;
  for (int d = 0; d < dim; ++d) {
    double r = prd[d] / dx;
    if (std::pow(((double)((int)(r + 0.5)) - r), 2.0) > SMALL)
      return fail(nullptr, LBMPSM_ERR_ARG,
                  "Illegal cell width. Division of domain length and cell width has to give an integer.");
  }
  // src/fix_LBM_PSM.cpp:142-154
  int n[3] = {0, 0, 0};
  for (int d = 0; d < dim; ++d) {
    n[d] = (int)(prd[d] / dx + 1.5);
    if (periodic[d]) n[d] -= 1;
  }
  cfg->dim = dim;
  cfg->nx = n[0];
  cfg->ny = n[1];
  cfg->nz = dim == 3 ? n[2] : 1;
  for (int d = 0; d < 3; ++d) cfg->periodic[d] = periodic[d] ? 1 : 0;
  if (dim == 2) cfg->periodic[2] = 1;
  cfg->tau = tau;
  // get_volume_force_lb, src/LBM_PSM_unitConversion.cpp:81-88
  for (int d = 0; d < 3; ++d) cfg->F_lbm[d] = Fext[d] / volumeForceFactor;
  cfg->origin[0] = boxlo[0];
  cfg->origin[1] = boxlo[1];
  cfg->origin[2] = dim == 2 ? 0.0 : boxlo[2];
  cfg->dx = dx;
  cfg->Uc = Uc;
  cfg->u_lb = u_lb;
  cfg->tc = tc;
  cfg->dt_d = dt_d;
  cfg->force_factor = forceFactor;
  cfg->torque_factor = torqueFactor;
  cfg->storage = LBMPSM_STORE_F64;
  cfg->device = 0;
  return LBMPSM_OK;
}

int lbmpsm_create(const lbmpsm_config* cfg, lbmpsm_ctx** out) {
  if (!cfg || !out) return fail(nullptr, LBMPSM_ERR_ARG, "null argument");
  *out = nullptr;
  if (cfg->dim != 2 && cfg->dim != 3) return fail(nullptr, LBMPSM_ERR_ARG, "dim must be 2 or 3");
  if (cfg->nx < 2 || cfg->ny < 2 || (cfg->dim == 3 && cfg->nz < 2))
    return fail(nullptr, LBMPSM_ERR_ARG, "need at least 2 nodes per axis");
  if (!(cfg->tau > 0.5)) return fail(nullptr, LBMPSM_ERR_ARG, "tau must be > 0.5");
  if (cfg->storage < LBMPSM_STORE_F64 || cfg->storage > LBMPSM_STORE_F16)
    return fail(nullptr, LBMPSM_ERR_ARG, "unknown storage form");
  if (!(cfg->dx > 0.0) || !(cfg->Uc != 0.0))
    return fail(nullptr, LBMPSM_ERR_ARG, "unit conversion factors not set");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
    return fail(nullptr, LBMPSM_ERR_NODEVICE, "no HIP device available");
  if (cfg->device < 0 || cfg->device >= ndev)
    return fail(nullptr, LBMPSM_ERR_ARG, "device ordinal out of range");

  lbmpsm_ctx* c = new (std::nothrow) lbmpsm_ctx();
  if (!c) return fail(nullptr, LBMPSM_ERR_NOMEM, "host allocation failed");
  c->cfg = *cfg;
  c->device = cfg->device;
  if (hipSetDevice(c->device) != hipSuccess) {
    delete c;
    return fail(nullptr, LBMPSM_ERR_HIP, "hipSetDevice failed");
  }
  if (noalloc_setup(c->device) != LBMPSM_OK) {
    delete c;
    return fail(nullptr, LBMPSM_ERR_HIP, "hipGetSymbolAddress of the static scratch failed");
  }
  Geom& g = c->g;
  g.dim = cfg->dim;
  g.q = cfg->dim == 3 ? 19 : 9;
  g.nx = cfg->nx;
  g.ny = cfg->ny;
  g.nz = cfg->dim == 3 ? cfg->nz : 1;
  const int glob[3] = {g.nx, g.ny, g.nz};
  int s[3];
  for (int d = 0; d < 3; ++d) {
    const bool active = d < cfg->dim;
    const bool per = cfg->periodic[d] != 0;
    g.per[d] = (!active || per) ? 1 : 0;
    if (!active) {
      s[d] = 1;
      g.off[d] = 0;
    } else if (per) {
      s[d] = glob[d];
      g.off[d] = 1;
    } else {
      s[d] = glob[d] + 2;
      g.off[d] = 0;
    }
  }
  g.sx = s[0];
  g.sy = s[1];
  g.sz = s[2];
  g.nodes = (int64_t)s[0] * s[1] * s[2];
  g.plane = (g.nodes + 63) / 64 * 64;
  if ((int64_t)g.plane >= (int64_t)1 << 31) {
    delete c;
    return fail(nullptr, LBMPSM_ERR_LIMIT, "lattice exceeds 2^31 stored nodes");
  }
  c->elem = cfg->storage == LBMPSM_STORE_F64 ? 8 : (cfg->storage == LBMPSM_STORE_F32 ? 4 : 2);
  if (cfg->storage == LBMPSM_STORE_F16) {
    c->scale = 4096.0f;
    c->inv_scale = 1.0f / 4096.0f;
  }
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return fail(nullptr, LBMPSM_ERR_HIP, "hipStreamCreate failed");
  }
  if (hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_sf, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_mapdone, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_collide, hipEventDisableTiming) != hipSuccess) {
    lbmpsm_destroy(c);
    return fail(nullptr, LBMPSM_ERR_HIP, "hipStreamCreate / hipEventCreate failed");
  }
  int rc = 0;
  const size_t P = (size_t)g.plane;
  rc |= dev_alloc(c, (char**)&c->f[0], P * g.q * c->elem);
  rc |= dev_alloc(c, (char**)&c->f[1], P * g.q * c->elem);
  rc |= dev_alloc(c, &c->slots.tag, 2 * P);
  rc |= dev_alloc(c, &c->slots.sf, 2 * P);
  rc |= dev_alloc(c, &c->slots.us, 6 * P);
  rc |= dev_alloc(c, &c->slots.Fh, 6 * P);
  rc |= dev_alloc(c, &c->slots.mask, P);
  if (c->cfg.storage != LBMPSM_STORE_F64) {
    c->slots.sff = reinterpret_cast<float*>(c->slots.sf);
    c->slots.usf = reinterpret_cast<float*>(c->slots.us);
    c->slots.Fhf = reinterpret_cast<float*>(c->slots.Fh);
  }
  rc |= dev_alloc(c, &c->d_cb, 19 + 10);  // the doubles, then the same as floats
  rc |= dev_alloc(c, &c->d_faces, kMaxFaces);
  // node bricks for the particle map: 256 nodes each
  BrickDev& B = c->br;
  if (g.dim == 3) {
    B.bi = kBrick3[0]; B.bj = kBrick3[1]; B.bk = kBrick3[2];
  } else {
    B.bi = kBrick2[0]; B.bj = kBrick2[1]; B.bk = kBrick2[2];
  }
  B.nbi = (g.sx + B.bi - 1) / B.bi;
  B.nbj = (g.sy + B.bj - 1) / B.bj;
  B.nbk = (g.sz + B.bk - 1) / B.bk;
  B.nbricks = (int64_t)B.nbi * B.nbj * B.nbk;
  rc |= dev_alloc(c, &B.count, B.nbricks);
  rc |= dev_alloc(c, &B.list, (size_t)B.nbricks * kBrickCap);
  rc |= dev_alloc(c, &B.slow_list, (size_t)B.nbricks);
  rc |= dev_alloc(c, &B.slow_count, (size_t)1);
  if (rc) {
    std::string m = c->err;
    lbmpsm_destroy(c);
    return fail(nullptr, LBMPSM_ERR_NOMEM, m);
  }
  if (hipEventCreateWithFlags(&c->stage_done, hipEventDisableTiming) != hipSuccess) {
    lbmpsm_destroy(c);
    return fail(nullptr, LBMPSM_ERR_HIP, "hipEventCreate failed");
  }
  if (launch_fill_int(B.count, B.nbricks, 0, c->stream) != hipSuccess) {
    lbmpsm_destroy(c);
    return fail(nullptr, LBMPSM_ERR_HIP, "brick count reset failed");
  }
  *out = c;
  // reference start: rho = 1, u = 0 (src/fix_LBM_PSM.cpp:168)
  rc = lbmpsm_init_equilibrium(c, 1.0, 0.0, 0.0, 0.0);
  if (rc) {
    g_create_error = c->err;
    lbmpsm_destroy(c);
    *out = nullptr;
    return rc;
  }
  return LBMPSM_OK;
}

int lbmpsm_destroy(lbmpsm_ctx* c) {
  if (!c) return LBMPSM_OK;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->stream2) (void)hipStreamSynchronize(c->stream2);
  dev_free(c->f[0]);
  dev_free(c->f[1]);
  dev_free(c->slots.tag);
  dev_free(c->slots.sf);
  dev_free(c->slots.us);
  dev_free(c->slots.Fh);
  dev_free(c->slots.mask);
  dev_free(c->rho);
  dev_free(c->u);
  ParticleDev& P = c->part;
  void* pp[] = {P.raw, P.xl, P.kin, P.zone, P.rzone, P.out, P.hpart,
                P.raw_alt, P.xl_alt, P.kin_alt, P.zone_alt, P.rzone_alt};
  for (void* p : pp) dev_free(p);
  BrickDev& B = c->br;
  void* bb[] = {B.count, B.list, B.slow_list, B.slow_count};
  for (void* p : bb) dev_free(p);
  dev_free(c->d_scratch);
  dev_free(c->sfcode);
  dev_free(c->d_cb);
  dev_free(c->d_faces);
  {
    auto& A = c->adj;
    void* aa[] = {A.ckpt, A.seg, A.lam, A.acc, A.seed, A.writer, A.bar, A.eout, A.partials,
                  A.Fh_save, A.mask_save};
    for (void* p : aa) dev_free(p);
  }
  if (c->h_stage) (void)hipHostFree(c->h_stage);
  if (c->h_out) (void)hipHostFree(c->h_out);
  for (auto& p : c->pending) {
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  for (auto e : c->free_events) (void)hipEventDestroy(e);
  if (c->stage_done) (void)hipEventDestroy(c->stage_done);
  if (c->ev_sf) (void)hipEventDestroy(c->ev_sf);
  if (c->ev_mapdone) (void)hipEventDestroy(c->ev_mapdone);
  if (c->ev_collide) (void)hipEventDestroy(c->ev_collide);
  if (c->stream2) (void)hipStreamDestroy(c->stream2);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return LBMPSM_OK;
}

int lbmpsm_get_info(const lbmpsm_ctx* c, lbmpsm_info* info) {
  if (!c || !info) return LBMPSM_ERR_ARG;
  info->dim = c->g.dim;
  info->q = c->g.q;
  info->nx = c->g.nx;
  info->ny = c->g.ny;
  info->nz = c->g.nz;
  info->sx = c->g.sx;
  info->sy = c->g.sy;
  info->sz = c->g.sz;
  for (int d = 0; d < 3; ++d) info->ref_offset[d] = c->g.off[d];
  info->stored_nodes = c->g.nodes;
  info->storage = c->cfg.storage;
  info->device_bytes = c->device_bytes;
  return LBMPSM_OK;
}

int lbmpsm_init_equilibrium(lbmpsm_ctx* c, double rho, double ux, double uy, double uz) {
  if (!c) return LBMPSM_ERR_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  const int Q = c->g.q;
  for (int q = 0; q < Q; ++q) {
    c->cb_raw[q] = feq_ref(Q, q, rho, ux, uy, uz);
    c->cb_store[q] = c->cfg.storage == LBMPSM_STORE_F64 ? c->cb_raw[q]
                                                        : c->cb_raw[q] - wq(Q, q) * c->rho0;
  }
  c->cur = 0;
  float cbf[19];
  for (int q = 0; q < 19; ++q) cbf[q] = (float)c->cb_store[q];
  HIPCHK(c, hipMemcpyAsync(c->d_cb, c->cb_store, sizeof(c->cb_store), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_cb + 19, cbf, sizeof(cbf), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, launch_fill_planes(c->cfg.storage, c->f[0], c->g, c->cb_store, c->rho0, c->scale, c->stream));
  HIPCHK(c, launch_slots_reset(c->g, c->slots, c->stream));
  c->poststream = 1;
  c->hydro_ready = 0;
  c->have_macros = 0;
  return LBMPSM_OK;
}

int lbmpsm_set_bc(lbmpsm_ctx* c, int keyword, double u_infty, double rho_outlet) {
  const int one[3] = {1, 1, 1};
  return lbmpsm_set_bc_ranks(c, keyword, u_infty, rho_outlet, one);
}

int lbmpsm_set_bc_ranks(lbmpsm_ctx* c, int keyword, double u_infty, double rho_outlet,
                        const int procgrid[3]) {
  if (!c || !procgrid) return LBMPSM_ERR_ARG;
  if (keyword < LBMPSM_BC_NONE || keyword > LBMPSM_BC_CLOSEDBOX)
    return fail(c, LBMPSM_ERR_ARG, "Illegal fix lbm-psm-bc command");
  const int nglob[3] = {c->g.nx, c->g.ny, c->g.nz};
  int64_t segs = 1;
  for (int d = 0; d < 3; ++d) {
    if (procgrid[d] < 1 || procgrid[d] > nglob[d])
      return fail(c, LBMPSM_ERR_ARG, "rank grid must have 1..n ranks per axis");
    segs = std::max<int64_t>(segs, (int64_t)procgrid[d] * procgrid[(d + 1) % 3]);
  }
  if (c->g.dim == 2 && procgrid[2] != 1)
    return fail(c, LBMPSM_ERR_ARG, "a 2D lattice has one rank along z");
  if (6 * segs > kMaxFaces)
    return fail(c, LBMPSM_ERR_LIMIT, "rank grid gives more face segments than the table holds");
  HIPCHK(c, hipSetDevice(c->device));
  c->bc_keyword = keyword;
  for (int d = 0; d < 3; ++d) c->procgrid[d] = procgrid[d];
  build_faces(c, keyword, u_infty, rho_outlet);
  if (c->nfaces)
    HIPCHK(c, hipMemcpyAsync(c->d_faces, c->faces, c->nfaces * sizeof(Face), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return LBMPSM_OK;
}

int lbmpsm_set_particles(lbmpsm_ctx* c, int n, const int* tag, const double* x, const double* v,
                         const double* omega, const double* radius) {
  if (!c) return LBMPSM_ERR_ARG;
  if (n < 0) return fail(c, LBMPSM_ERR_ARG, "negative particle count");
  if (n > 0 && (!tag || !x || !v || !omega || !radius)) return fail(c, LBMPSM_ERR_ARG, "null particle array");
  HIPCHK(c, hipSetDevice(c->device));
  c->hydro_ready = 0;
  c->part.n = n;
  if (n == 0) return LBMPSM_OK;
  int rc = ensure_particles(c, n);
  if (rc) return rc;
  double rmax = 0.0;
  for (int p = 0; p < n; ++p) {
    if (!(std::fabs(radius[p]) < 1e300)) return fail(c, LBMPSM_ERR_ARG, "radius is not finite");
    rmax = std::max(rmax, std::fabs(radius[p]));
  }
  rc = size_codes(c, n, rmax);
  if (rc) return rc;
  rc = next_list(c);
  if (rc) return rc;
  // stage into pinned memory (after the previous upload has left it), one copy
  HIPCHK(c, hipEventSynchronize(c->stage_done));
  char* h = (char*)c->h_stage;
  double* hd = (double*)h;
  std::memcpy(hd, x, 3 * (size_t)n * sizeof(double));
  std::memcpy(hd + 3 * (size_t)n, v, 3 * (size_t)n * sizeof(double));
  std::memcpy(hd + 6 * (size_t)n, omega, 3 * (size_t)n * sizeof(double));
  std::memcpy(hd + 9 * (size_t)n, radius, (size_t)n * sizeof(double));
  std::memcpy(h + 80 * (size_t)n, tag, (size_t)n * sizeof(int));
  HIPCHK(c, hipMemcpyAsync(c->part.raw, h, 84 * (size_t)n, hipMemcpyHostToDevice, c->stream2));
  HIPCHK(c, hipEventRecord(c->stage_done, c->stream2));
  return run_map(c);
}

int lbmpsm_set_particles_device(lbmpsm_ctx* c, int n, const int* tag, const double* x,
                                const double* v, const double* omega, const double* radius,
                                double max_radius) {
  if (!c) return LBMPSM_ERR_ARG;
  if (n < 0) return fail(c, LBMPSM_ERR_ARG, "negative particle count");
  if (n > 0 && (!tag || !x || !v || !omega || !radius)) return fail(c, LBMPSM_ERR_ARG, "null particle array");
  if (!(max_radius >= 0.0 && max_radius < 1e300))
    return fail(c, LBMPSM_ERR_ARG, "max_radius must be finite and >= 0");
  HIPCHK(c, hipSetDevice(c->device));
  c->hydro_ready = 0;
  c->part.n = n;
  if (n == 0) return LBMPSM_OK;
  int rc = ensure_particles(c, n);
  if (rc) return rc;
  rc = size_codes(c, n, max_radius);
  if (rc) return rc;
  rc = next_list(c);
  if (rc) return rc;
  const DevList dev{tag, x, v, omega, radius};
  return run_map(c, &dev);
}

int lbmpsm_step(lbmpsm_ctx* c, int flags) {
  if (!c) return LBMPSM_ERR_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  const int store = (flags & LBMPSM_STEP_STORE_MACROS) ? 1 : 0;
  if (store && !c->rho) {
    int rc = dev_alloc(c, &c->rho, (size_t)c->g.plane);
    rc |= dev_alloc(c, &c->u, 3 * (size_t)c->g.plane);
    if (rc) return LBMPSM_ERR_NOMEM;
  }
  return step_launch(c, store, c->part.hpart, c->part.out, true);
}

int lbmpsm_step_noalloc(lbmpsm_ctx* c, int flags, double* out) {
  if (!c) return LBMPSM_ERR_ARG;
  const int n = c->part.n;
  if (n > 0 && !out) return fail(c, LBMPSM_ERR_ARG, "null output array");
  HIPCHK(c, hipSetDevice(c->device));
  const int store = (flags & LBMPSM_STEP_STORE_MACROS) ? 1 : 0;
  if (store && !c->rho)
    return fail(c, LBMPSM_ERR_STATE,
                "LBMPSM_STEP_STORE_MACROS needs the macro arrays of an earlier lbmpsm_step "
                "with that flag (this entry point allocates nothing)");
  if ((int64_t)n * kHydroGroups * 12 > kNoAllocScratchDoubles)
    return fail(c, LBMPSM_ERR_LIMIT,
                "particle list longer than the static scratch holds (" +
                    std::to_string(kNoAllocScratchDoubles / (kHydroGroups * 12)) + " entries)");
  double* scr = noalloc_scratch(c->device);
  if (!scr) return fail(c, LBMPSM_ERR_STATE, "static scratch address not set up on this device");
  return step_launch(c, store, scr, out, false);
}

int lbmpsm_get_hydro(lbmpsm_ctx* c, int n, double* F, double* T, double* S) {
  if (!c) return LBMPSM_ERR_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  if (!c->hydro_ready) return fail(c, LBMPSM_ERR_STATE, "no step has run since the particles were set");
  if (n != c->hydro_n) return fail(c, LBMPSM_ERR_ARG, "n differs from the particle count of the last step");
  if (n == 0) return LBMPSM_OK;
  int rc = ensure_host_out(c, (size_t)n * 12 * sizeof(double));
  if (rc) return rc;
  HIPCHK(c, hipMemcpyAsync(c->h_out, c->part.out, (size_t)n * 12 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int p = 0; p < n; ++p) {
    const double* o = c->h_out + 12 * (size_t)p;
    if (F) for (int d = 0; d < 3; ++d) F[3 * p + d] = o[d];
    if (T) for (int d = 0; d < 3; ++d) T[3 * p + d] = o[3 + d];
    if (S) for (int d = 0; d < 6; ++d) S[6 * p + d] = o[6 + d];
  }
  return LBMPSM_OK;
}

int lbmpsm_get_populations(lbmpsm_ctx* c, double* f) {
  if (!c || !f) return LBMPSM_ERR_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  const size_t count = (size_t)c->g.nx * c->g.ny * c->g.nz * c->g.q;
  int rc = ensure_scratch(c, count * sizeof(double));
  if (rc) return rc;
  CollideArgs a = make_args(c, 0);
  HIPCHK(c, launch_finish(c->cfg.storage, a, c->d_scratch, c->stream));
  HIPCHK(c, hipMemcpyAsync(f, c->d_scratch, count * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return LBMPSM_OK;
}

int lbmpsm_set_populations(lbmpsm_ctx* c, const double* f) {
  if (!c || !f) return LBMPSM_ERR_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  const size_t count = (size_t)c->g.nx * c->g.ny * c->g.nz * c->g.q;
  int rc = ensure_scratch(c, count * sizeof(double));
  if (rc) return rc;
  HIPCHK(c, hipMemcpyAsync(c->d_scratch, f, count * sizeof(double), hipMemcpyHostToDevice, c->stream));
  // halo nodes of non-periodic axes restart from the initial equilibrium
  HIPCHK(c, launch_fill_planes(c->cfg.storage, c->f[c->cur], c->g, c->cb_store, c->rho0, c->scale, c->stream));
  HIPCHK(c, launch_set_populations(c->cfg.storage, c->f[c->cur], c->g, c->d_scratch, c->rho0, c->scale, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->poststream = 1;
  return LBMPSM_OK;
}

static size_t ref_count(const lbmpsm_ctx* c) {
  const Geom& g = c->g;
  return (size_t)(g.nx + 2) * (g.ny + 2) * (g.dim == 3 ? g.nz + 2 : 1) * g.q;
}

int lbmpsm_get_populations_ref(lbmpsm_ctx* c, double* f) {
  if (!c || !f) return LBMPSM_ERR_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  const size_t count = ref_count(c);
  int rc = ensure_scratch(c, count * sizeof(double));
  if (rc) return rc;
  CollideArgs a = make_args(c, 0);
  HIPCHK(c, launch_finish_ref(c->cfg.storage, a, c->d_scratch, c->stream));
  HIPCHK(c, hipMemcpyAsync(f, c->d_scratch, count * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return LBMPSM_OK;
}

int lbmpsm_set_populations_ref(lbmpsm_ctx* c, const double* f) {
  if (!c || !f) return LBMPSM_ERR_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  const size_t count = ref_count(c);
  int rc = ensure_scratch(c, count * sizeof(double));
  if (rc) return rc;
  HIPCHK(c, hipMemcpyAsync(c->d_scratch, f, count * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, launch_set_populations_ref(c->cfg.storage, c->f[c->cur], c->g, c->d_scratch, c->rho0, c->scale, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->poststream = 1;
  return LBMPSM_OK;
}

int lbmpsm_get_macros(lbmpsm_ctx* c, double* rho, double* u) {
  if (!c) return LBMPSM_ERR_ARG;
  if (!c->have_macros) return fail(c, LBMPSM_ERR_STATE, "no step stored macroscopic fields");
  HIPCHK(c, hipSetDevice(c->device));
  const size_t nn = (size_t)c->g.nx * c->g.ny * c->g.nz;
  int rc = ensure_scratch(c, nn * 4 * sizeof(double));
  if (rc) return rc;
  double* dr = c->d_scratch;
  double* du = c->d_scratch + nn;
  HIPCHK(c, launch_macros_export(c->g, c->rho, c->u, dr, du, c->stream));
  if (rho) HIPCHK(c, hipMemcpyAsync(rho, dr, nn * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (u) HIPCHK(c, hipMemcpyAsync(u, du, 3 * nn * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return LBMPSM_OK;
}

int lbmpsm_get_slots(lbmpsm_ctx* c, int* tag, double* sf, double* us, double* F) {
  if (!c) return LBMPSM_ERR_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  const size_t nn = (size_t)c->g.nx * c->g.ny * c->g.nz;
  // scratch: tag (2 ints per node, stored as doubles' space) + 14 doubles
  int rc = ensure_scratch(c, nn * (2 * sizeof(int) + 14 * sizeof(double)) + 64);
  if (rc) return rc;
  double* dsf = c->d_scratch;
  double* dus = dsf + 2 * nn;
  double* dF = dus + 6 * nn;
  int* dtag = (int*)(dF + 6 * nn);
  HIPCHK(c, launch_slots_export(c->g, c->slots, dtag, dsf, dus, dF, c->stream));
  if (tag) HIPCHK(c, hipMemcpyAsync(tag, dtag, 2 * nn * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  if (sf) HIPCHK(c, hipMemcpyAsync(sf, dsf, 2 * nn * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (us) HIPCHK(c, hipMemcpyAsync(us, dus, 6 * nn * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (F) HIPCHK(c, hipMemcpyAsync(F, dF, 6 * nn * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return LBMPSM_OK;
}

int lbmpsm_mass(lbmpsm_ctx* c, double* mass) {
  if (!c || !mass) return LBMPSM_ERR_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  const int nb = (int)std::min<int64_t>(1024, (c->g.nodes + 255) / 256);
  int rc = ensure_scratch(c, (size_t)nb * sizeof(double));
  if (rc) return rc;
  CollideArgs a = make_args(c, 0);
  HIPCHK(c, launch_mass(c->cfg.storage, a, c->d_scratch, nb, c->stream));
  std::vector<double> part(nb);
  HIPCHK(c, hipMemcpyAsync(part.data(), c->d_scratch, nb * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  double m = 0.0;
  for (int i = 0; i < nb; ++i) m += part[i];
  *mass = m;
  return LBMPSM_OK;
}

int lbmpsm_sync(lbmpsm_ctx* c) {
  if (!c) return LBMPSM_ERR_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return LBMPSM_OK;
}

int lbmpsm_profile_enable(lbmpsm_ctx* c, int enable) {
  if (!c) return LBMPSM_ERR_ARG;
  if (enable < 0 || enable > 2) return fail(c, LBMPSM_ERR_ARG, "profile mode must be 0, 1 or 2");
  c->prof = enable;
  return LBMPSM_OK;
}

int lbmpsm_profile_read(lbmpsm_ctx* c, const char* kernel, double* total_ms, int64_t* count) {
  if (!c || !kernel) return LBMPSM_ERR_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  int rc = prof_collect(c);
  if (rc) return rc;
  auto it = c->totals.find(kernel);
  if (total_ms) *total_ms = it == c->totals.end() ? 0.0 : it->second.first;
  if (count) *count = it == c->totals.end() ? 0 : it->second.second;
  return LBMPSM_OK;
}

int lbmpsm_profile_reset(lbmpsm_ctx* c) {
  if (!c) return LBMPSM_ERR_ARG;
  int rc = prof_collect(c);
  c->totals.clear();
  return rc;
}


}  // extern "C"

// ---------------------------------------------------------------------------
// Adjoint of a run with frozen particles (FP64, fully periodic, no faces).
// ---------------------------------------------------------------------------
namespace {

template <typename T>
int adj_grow(lbmpsm_ctx* c, T** p, size_t& have, size_t need) {
  if (*p && have >= need) return LBMPSM_OK;
  if (*p) {
    HIPCHK(c, hipStreamSynchronize(c->stream));
    dev_free(*p);
    *p = nullptr;
    c->device_bytes -= (int64_t)(have * sizeof(T));
    have = 0;
  }
  int rc = dev_alloc(c, p, need);
  if (rc) return rc;
  have = need;
  return LBMPSM_OK;
}

int adj_supported(lbmpsm_ctx* c) {
  if (c->cfg.storage != LBMPSM_STORE_F64)
    return fail(c, LBMPSM_ERR_STATE, "the adjoint runs on FP64 storage only");
  if (c->nfaces)
    return fail(c, LBMPSM_ERR_STATE, "the adjoint supports lattices without Zou-He faces");
  for (int d = 0; d < c->g.dim; ++d)
    if (!c->g.per[d])
      return fail(c, LBMPSM_ERR_STATE, "the adjoint supports fully periodic lattices only");
  return LBMPSM_OK;
}

}  // namespace

extern "C" {

int lbmpsm_adjoint_forward(lbmpsm_ctx* c, int steps, int checkpoint_every, int n,
                           const int* tag, const double* x, const double* v,
                           const double* omega, const double* radius, double* F, double* T,
                           double* S) {
  if (!c) return LBMPSM_ERR_ARG;
  if (steps < 1 || checkpoint_every < 1)
    return fail(c, LBMPSM_ERR_ARG, "steps and checkpoint_every must be >= 1");
  int rc = adj_supported(c);
  if (rc) return rc;
  HIPCHK(c, hipSetDevice(c->device));
  auto& A = c->adj;
  A.ready = 0;
  const size_t S1 = (size_t)c->g.q * (size_t)c->g.plane;
  const int ncp = (steps + checkpoint_every - 1) / checkpoint_every;
  rc = adj_grow(c, &A.ckpt, A.ckpt_n, (size_t)ncp * S1);
  if (rc) return rc;
  A.post.assign(ncp, 0);
  HIPCHK(c, hipMemcpyAsync(A.ckpt, c->f[c->cur], S1 * sizeof(double), hipMemcpyDeviceToDevice,
                           c->stream));
  A.post[0] = c->poststream;
  for (int t = 1; t <= steps; ++t) {
    rc = lbmpsm_set_particles(c, n, tag, x, v, omega, radius);
    if (rc) return rc;
    rc = lbmpsm_step(c, 0);
    if (rc) return rc;
    if (t < steps && t % checkpoint_every == 0) {
      const int k = t / checkpoint_every;
      HIPCHK(c, hipMemcpyAsync(A.ckpt + (size_t)k * S1, c->f[c->cur], S1 * sizeof(double),
                               hipMemcpyDeviceToDevice, c->stream));
      A.post[k] = c->poststream;
    }
  }
  if (F || T || S) {
    rc = lbmpsm_get_hydro(c, n, F, T, S);
    if (rc) return rc;
  }
  A.steps = steps;
  A.every = checkpoint_every;
  A.ncp = ncp;
  A.n = n;
  A.ready = 1;
  return LBMPSM_OK;
}

int lbmpsm_adjoint_backward(lbmpsm_ctx* c, int n, const double* Fbar, const double* Tbar,
                            const double* Sbar, double* grad_F_lbm, double* grad_tau,
                            double* grad_v, double* grad_omega) {
  if (!c) return LBMPSM_ERR_ARG;
  auto& A = c->adj;
  if (!A.ready) return fail(c, LBMPSM_ERR_STATE, "lbmpsm_adjoint_forward has not run");
  if (n != A.n || n != c->part.n)
    return fail(c, LBMPSM_ERR_ARG, "n differs from the particle count of the forward run");
  int rc = adj_supported(c);
  if (rc) return rc;
  HIPCHK(c, hipSetDevice(c->device));
  const Geom& g = c->g;
  const size_t P = (size_t)g.plane;
  const size_t S1 = (size_t)g.q * P;
  rc = adj_grow(c, &A.seg, A.seg_n, (size_t)std::max(A.every - 1, 1) * S1);
  if (rc) return rc;
  if (!A.lam) {  // plane-sized buffers: allocated once per context
    rc = dev_alloc(c, &A.lam, 2 * S1);
    rc |= dev_alloc(c, &A.acc, 10 * P);
    rc |= dev_alloc(c, &A.seed, 6 * P);
    rc |= dev_alloc(c, &A.writer, 2 * P);
    rc |= dev_alloc(c, &A.partials, 1024 * 4);
    rc |= dev_alloc(c, &A.Fh_save, 6 * P);
    rc |= dev_alloc(c, &A.mask_save, P);
    if (rc) return LBMPSM_ERR_NOMEM;
  }
  if (A.ent_n < (size_t)std::max(n, 1)) {  // per-entry buffers grow with n
    HIPCHK(c, hipStreamSynchronize(c->stream));
    dev_free(A.bar);
    dev_free(A.eout);
    A.bar = A.eout = nullptr;
    c->device_bytes -= (int64_t)(A.ent_n * 18 * sizeof(double));
    A.ent_n = 0;
    rc = dev_alloc(c, &A.bar, (size_t)std::max(n, 1) * 12);
    rc |= dev_alloc(c, &A.eout, (size_t)std::max(n, 1) * 6);
    if (rc) return LBMPSM_ERR_NOMEM;
    A.ent_n = (size_t)std::max(n, 1);
  }

  // per-entry loss gradients
  std::vector<double> hb((size_t)std::max(n, 1) * 12, 0.0);
  for (int e = 0; e < n; ++e)
    for (int d = 0; d < 3; ++d) {
      if (Fbar) hb[12 * e + d] = Fbar[3 * e + d];
      if (Tbar) hb[12 * e + 3 + d] = Tbar[3 * e + d];
    }
  if (Sbar)
    for (int e = 0; e < n; ++e)
      for (int c6 = 0; c6 < 6; ++c6) hb[12 * e + 6 + c6] = Sbar[6 * e + c6];
  HIPCHK(c, hipMemcpyAsync(A.bar, hb.data(), hb.size() * sizeof(double), hipMemcpyHostToDevice,
                           c->stream));
  AdjSeedArgs sa{};
  sa.g = g;
  sa.n = n;
  sa.tag = c->part.tag(c->part.n);
  sa.xl = c->part.xl;
  sa.zone = c->part.zone;
  sa.rzone = c->part.rzone;
  sa.slot_tag = c->slots.tag;
  sa.mask = c->slots.mask;
  sa.bar = A.bar;
  sa.seed = A.seed;
  sa.writer = A.writer;
  sa.forceFactor = c->cfg.force_factor;
  sa.torqueFactor = c->cfg.torque_factor;
  HIPCHK(c, launch_adj_seed(sa, c->stream));
  HIPCHK(c, launch_adj_writer(sa, c->stream));
  HIPCHK(c, hipMemsetAsync(A.acc, 0, 10 * P * sizeof(double), c->stream));

  // The recomputing collides below accumulate node forces into the slots and
  // clear fresh bits, as any collide does; the forward's final slot forces and
  // masks are saved here and put back at the end, so that the context leaves
  // backward in the state the forward left it.
  HIPCHK(c, hipMemcpyAsync(A.Fh_save, c->slots.Fh, 6 * P * sizeof(double),
                           hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(A.mask_save, c->slots.mask, P, hipMemcpyDeviceToDevice, c->stream));
  CollideArgs base = make_args(c, 0);
  AdjArgs aa{};
  aa.g = g;
  aa.mask = c->slots.mask;
  aa.sf = c->slots.sf;
  aa.us = c->slots.us;
  aa.acc = A.acc;
  aa.omega = base.omega;
  aa.invCs2 = base.invCs2;
  aa.invCs4 = base.invCs4;
  for (int d = 0; d < 3; ++d) aa.F[d] = base.F[d];
  double* lam[2] = {A.lam, A.lam + S1};
  int lc = 0;
  bool have_lam = false;
  for (int sg = A.ncp - 1; sg >= 0; --sg) {
    const int t0 = sg * A.every;
    const int t1 = std::min(t0 + A.every, A.steps);
    const int L = t1 - t0;
    // X[0] = checkpoint, X[i] = post-collision state after step t0 + i
    auto X = [&](int i) -> double* {
      return i == 0 ? A.ckpt + (size_t)sg * S1 : A.seg + (size_t)(i - 1) * S1;
    };
    for (int i = 1; i < L; ++i) {
      CollideArgs a = base;
      a.fin = X(i - 1);
      a.fout = X(i);
      a.poststream = i == 1 ? A.post[sg] : 0;
      HIPCHK(c, launch_collide(LBMPSM_STORE_F64, a, c->stream));
    }
    for (int t = t1; t > t0; --t) {
      const int i = t - t0;
      aa.hprev = X(i - 1);
      aa.inplace = i == 1 ? A.post[sg] : 0;
      aa.lam_in = have_lam ? lam[lc] : nullptr;
      aa.lam_out = lam[1 - lc];
      aa.seed = t == A.steps ? A.seed : nullptr;
      HIPCHK(c, launch_adj_step(aa, c->stream));
      lc = 1 - lc;
      have_lam = true;
    }
  }
  HIPCHK(c, hipMemcpyAsync(c->slots.Fh, A.Fh_save, 6 * P * sizeof(double),
                           hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->slots.mask, A.mask_save, P, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, launch_adj_entry(sa, A.acc, A.eout, c->stream));
  const int nb = (int)std::min<int64_t>(1024, (g.nodes + 255) / 256);
  HIPCHK(c, launch_adj_param(g, A.acc, A.partials, nb, c->stream));
  std::vector<double> eo((size_t)std::max(n, 1) * 6), pp((size_t)nb * 4);
  if (n > 0)
    HIPCHK(c, hipMemcpyAsync(eo.data(), A.eout, (size_t)n * 6 * sizeof(double),
                             hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(pp.data(), A.partials, pp.size() * sizeof(double),
                           hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  double gp[4] = {0.0, 0.0, 0.0, 0.0};
  for (int b = 0; b < nb; ++b)
    for (int k = 0; k < 4; ++k) gp[k] += pp[4 * b + k];
  // om = 1/tau; v_lb = v / Uc * u_lb; w_lb = w * tc * dt_d
  if (grad_tau) *grad_tau = gp[0] * (-1.0 / (c->cfg.tau * c->cfg.tau));
  if (grad_F_lbm)
    for (int d = 0; d < 3; ++d) grad_F_lbm[d] = gp[1 + d];
  const double sv = c->cfg.u_lb / c->cfg.Uc, sw = c->cfg.tc * c->cfg.dt_d;
  for (int e = 0; e < n; ++e)
    for (int d = 0; d < 3; ++d) {
      if (grad_v) grad_v[3 * e + d] = eo[6 * e + d] * sv;
      if (grad_omega) grad_omega[3 * e + d] = eo[6 * e + 3 + d] * sw;
    }
  return LBMPSM_OK;
}

}  // extern "C"
