"""ctypes wrapper of the CPU oracle (liblbmpsm_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg, always as the checker, never as the product.
The oracle restates the reference (timnajuch/LAMMPS-LBM-PSM src/) in C; see
lbmpsm_oracle.h for the function-to-reference-line map.

Arrays come back in the reference's own halo'd layouts; `interior_*`
helpers cut out the interior nodes in the order the GPU library's getters
use ([nx][ny][nz][...], nz = 1 in 2D).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblbmpsm_oracle.so")


class Units(C.Structure):
    _fields_ = [
        ("rhof", C.c_double), ("nu", C.c_double), ("lc", C.c_double), ("Re", C.c_double),
        ("tau", C.c_double), ("N", C.c_int), ("dim", C.c_int),
        ("Uc_d", C.c_double), ("Lc_d", C.c_double), ("cs", C.c_double), ("csPow2", C.c_double),
        ("nu_lb", C.c_double), ("dx", C.c_double), ("dx_d", C.c_double), ("Uc", C.c_double),
        ("u_lb", C.c_double), ("dt_d", C.c_double), ("tc", C.c_double),
        ("forceFactor", C.c_double), ("torqueFactor", C.c_double),
        ("volumeForceFactor", C.c_double),
    ]


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} not built (make -C oracle)")
    lib = C.CDLL(LIB_PATH)
    P = C.c_void_p
    dp = C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int)
    sig = {
        "lbo_units_init": (None, [C.POINTER(Units), C.c_double, C.c_double, C.c_double, C.c_double,
                                  C.c_int, C.c_double, C.c_int]),
        "lbo_create": (P, [C.c_int, C.c_int, C.c_int, C.c_int, ip, C.c_double, dp]),
        "lbo_destroy": (None, [P]),
        "lbo_dims": (None, [P, ip]),
        "lbo_initialise": (None, [P, C.c_double, C.c_double, C.c_double, C.c_double]),
        "lbo_set_particles": (None, [P, C.POINTER(Units), dp, C.c_int, ip, dp, dp, dp, dp]),
        "lbo_solid_fraction": (C.c_double, [P, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                            C.c_double, C.c_double]),
        "lbo_halo_exchange": (None, [P]),
        "lbo_collide_stream": (None, [P]),
        "lbo_apply_bc": (None, [P, C.c_int, C.c_double]),
        "lbo_zou_he_face": (C.c_int, [P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                      C.c_double, C.c_double, C.c_double, C.c_double]),
        "lbo_hydro": (None, [P, C.POINTER(Units), dp, C.c_int, dp, C.c_double, dp, dp, dp]),
        "lbo_fix_step": (None, [P, C.POINTER(Units), dp, C.c_int, ip, dp, dp, dp, dp, C.c_int,
                                dp, dp, dp]),
        "lbo_f": (dp, [P]),
        "lbo_rho": (dp, [P]),
        "lbo_u": (dp, [P]),
        "lbo_slot_tag": (ip, [P]),
        "lbo_slot_sf": (dp, [P]),
        "lbo_slot_u": (dp, [P]),
        "lbo_slot_F": (dp, [P]),
        "lbo_mass": (C.c_double, [P]),
        "lbo_set_rank": (None, [P, ip, ip]),
        "lbo_threads": (C.c_int, []),
        "lbo_set_threads": (None, [C.c_int]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


LIB = _load()


def _d(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(
        C.POINTER(C.c_double))


def units(rhof, nu, lc, Re, N, tau, dim) -> Units:
    u = Units()
    LIB.lbo_units_init(C.byref(u), rhof, nu, lc, Re, N, tau, dim)
    return u


def lattice_units(u_lb=0.01) -> Units:
    """dx = 1, velocities and forces unscaled (matches lattice_units_config):
    Uc = u_lb, so v_lb = v / Uc * u_lb = v; u_lb is also the Zou-He speed."""
    u = Units()
    u.dx = 1.0
    u.Uc = u_lb
    u.u_lb = u_lb
    u.tc = 1.0
    u.dt_d = 1.0
    u.forceFactor = 1.0
    u.torqueFactor = 1.0
    u.volumeForceFactor = 1.0
    return u


class Oracle:
    """One reference-layout lattice on the CPU."""

    def __init__(self, dim, n, periodic, tau, F_lbm=(0.0, 0.0, 0.0), unit=None,
                 origin=(0.0, 0.0, 0.0)):
        per = (C.c_int * 3)(*[int(p) for p in periodic])
        F = np.ascontiguousarray(F_lbm, dtype=np.float64)
        nz = n[2] if dim == 3 else 1
        self._h = LIB.lbo_create(dim, int(n[0]), int(n[1]), int(nz), per, tau, _d(F))
        if not self._h:
            raise ValueError("bad oracle lattice arguments")
        self._F = F
        dims = (C.c_int * 4)()
        LIB.lbo_dims(self._h, dims)
        self.nh = (dims[0], dims[1], dims[2])
        self.q = dims[3]
        self.dim = dim
        self.unit = unit if unit is not None else lattice_units()
        self.origin = np.ascontiguousarray(origin, dtype=np.float64)
        LIB.lbo_initialise(self._h, 1.0, 0.0, 0.0, 0.0)

    def close(self):
        if self._h:
            LIB.lbo_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def initialise(self, rho=1.0, u=(0.0, 0.0, 0.0)):
        LIB.lbo_initialise(self._h, rho, u[0], u[1], u[2])

    def fix_step(self, lst, bc_keyword=0):
        n = lst.n
        F = np.zeros((n, 3))
        T = np.zeros((n, 3))
        S = np.zeros((n, 6))
        tag = np.ascontiguousarray(lst.tag, dtype=np.int32)
        LIB.lbo_fix_step(self._h, C.byref(self.unit), _d(self.origin), n,
                         tag.ctypes.data_as(C.POINTER(C.c_int)), _d(lst.x), _d(lst.v),
                         _d(lst.omega), _d(lst.radius), int(bc_keyword), _d(F), _d(T), _d(S))
        return F, T, S

    def map(self, lst):
        """Only the particle map of a step (setParticlesOnLattice)."""
        tag = np.ascontiguousarray(lst.tag, dtype=np.int32)
        LIB.lbo_set_particles(self._h, C.byref(self.unit), _d(self.origin), lst.n,
                              tag.ctypes.data_as(C.POINTER(C.c_int)), _d(lst.x), _d(lst.v),
                              _d(lst.omega), _d(lst.radius))

    def set_rank(self, ntotal, pcoord):
        """Make this lattice one rank's sub-domain of a decomposed run
        (lbo_set_rank); its halos are then the caller's to fill."""
        LIB.lbo_set_rank(self._h, (C.c_int * 3)(*[int(v) for v in ntotal]),
                         (C.c_int * 3)(*[int(v) for v in pcoord]))

    def halo_exchange(self):
        LIB.lbo_halo_exchange(self._h)

    def collide_stream(self):
        LIB.lbo_collide_stream(self._h)

    def hydro(self, lst):
        """Force, torque and stresslet of every list entry over this
        lattice's interior nodes (calculateHydrodynamicInteractions)."""
        n = lst.n
        F, T, S = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 6))
        x = np.ascontiguousarray(lst.x, dtype=np.float64)
        for p in range(n):
            LIB.lbo_hydro(self._h, C.byref(self.unit), _d(self.origin), int(lst.tag[p]),
                          _d(x[p]), float(lst.radius[p]), _d(F[p]), _d(T[p]), _d(S[p]))
        return F, T, S

    def f_view(self):
        """The current (post-stream) populations with halos, in place (the
        buffer changes at every collide_stream)."""
        ptr = LIB.lbo_f(self._h)
        return np.ctypeslib.as_array(ptr, shape=(self.nodes_h * self.q,)).reshape(*self.nh,
                                                                                   self.q)

    def zou_he_face(self, kind, plane, a, b=(0, 0), u=(0.0, 0.0, 0.0), rho=1.0):
        """One face function on plane x [a0,a1] x [b0,b1] (reference halo'd
        indices) with the imposed velocity u or density rho."""
        rc = LIB.lbo_zou_he_face(self._h, int(kind), int(plane), int(a[0]), int(a[1]), int(b[0]),
                                 int(b[1]), float(u[0]), float(u[1]), float(u[2]), float(rho))
        if rc:
            raise ValueError(f"unknown face kind {kind}")

    def solid_fraction(self, i, j, k, xp, yp, zp, rp):
        return LIB.lbo_solid_fraction(self._h, i, j, k, xp, yp, zp, rp)

    # --- raw reference-layout views (copies) --------------------------------
    def _view(self, fn, count, dtype=np.float64):
        ptr = fn(self._h)
        return np.ctypeslib.as_array(ptr, shape=(count,)).astype(dtype, copy=True)

    @property
    def nodes_h(self):
        return self.nh[0] * self.nh[1] * self.nh[2]

    def f_halo(self):
        return self._view(LIB.lbo_f, self.nodes_h * self.q).reshape(*self.nh, self.q)

    def set_f_halo(self, f):
        ptr = LIB.lbo_f(self._h)
        arr = np.ctypeslib.as_array(ptr, shape=(self.nodes_h * self.q,))
        arr[:] = np.ascontiguousarray(f, dtype=np.float64).reshape(-1)

    def _interior(self, a):
        if self.dim == 3:
            return np.ascontiguousarray(a[1:-1, 1:-1, 1:-1])
        return np.ascontiguousarray(a[1:-1, 1:-1, :])

    def populations(self):
        return self._interior(self.f_halo())

    def macros(self):
        rho = self._view(LIB.lbo_rho, self.nodes_h).reshape(self.nh)
        u = self._view(LIB.lbo_u, self.nodes_h * 3).reshape(*self.nh, 3)
        return self._interior(rho), self._interior(u)

    def slots(self):
        tag = self._view(LIB.lbo_slot_tag, self.nodes_h * 2, np.int32).reshape(*self.nh, 2)
        sf = self._view(LIB.lbo_slot_sf, self.nodes_h * 2).reshape(*self.nh, 2)
        us = self._view(LIB.lbo_slot_u, self.nodes_h * 6).reshape(*self.nh, 6)
        F = self._view(LIB.lbo_slot_F, self.nodes_h * 6).reshape(*self.nh, 6)
        return self._interior(tag), self._interior(sf), self._interior(us), self._interior(F)

    def mass(self):
        return LIB.lbo_mass(self._h)


def threads():
    return LIB.lbo_threads()


def set_threads(n):
    LIB.lbo_set_threads(int(n))
