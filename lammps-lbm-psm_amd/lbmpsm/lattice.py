"""Python handle on one lbmpsm context (one lattice on one GPU).

Thin object layer over the C ABI; every method is one C call.  Arrays cross
the boundary as contiguous numpy float64/int32 arrays in the reference's
layouts (see include/lbmpsm.h).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N


def config_from_fix(dim, Nlc, lc, rho, nu, Re, tau=0.7, Fext=(0.0, 0.0, 0.0),
                    boxlo=(0.0, 0.0, 0.0), prd=(1.0, 1.0, 1.0), periodic=(1, 1, 1),
                    storage="f64", device=0) -> N.Config:
    """`fix lbm-psm` arguments + LAMMPS box -> lbmpsm_config (reference formulas)."""
    cfg = N.Config()
    F = np.ascontiguousarray(Fext, dtype=np.float64)
    lo = np.ascontiguousarray(boxlo, dtype=np.float64)
    pr = np.ascontiguousarray(prd, dtype=np.float64)
    per = np.ascontiguousarray(periodic, dtype=np.int32)
    N.check(N.LIB.lbmpsm_config_from_fix(C.byref(cfg), dim, Nlc, lc, rho, nu, Re, tau,
                                         N.dptr(F), N.dptr(lo), N.dptr(pr), N.iptr(per)))
    cfg.storage = N.STORAGE[storage] if isinstance(storage, str) else int(storage)
    cfg.device = device
    return cfg


def lattice_units_config(dim, n, periodic, tau, F_lbm=(0.0, 0.0, 0.0), storage="f64",
                         device=0, u_lb=0.01) -> N.Config:
    """Configuration in lattice units: dx = 1, origin 0, velocities, angular
    velocities, forces and torques unscaled (Uc = u_lb, tc = dt_d = 1,
    force and torque factors 1).  Particle positions are lattice coordinates
    with node i at x = i.  u_lb is the Zou-He speed lbmpsm_set_bc uses when
    the fix passes UnitConversion::get_u_lb()."""
    cfg = N.Config()
    cfg.dim = dim
    cfg.nx, cfg.ny = int(n[0]), int(n[1])
    cfg.nz = int(n[2]) if dim == 3 else 1
    for d in range(3):
        cfg.periodic[d] = int(periodic[d]) if d < dim else 1
    cfg.tau = tau
    for d in range(3):
        cfg.F_lbm[d] = F_lbm[d]
        cfg.origin[d] = 0.0
    cfg.dx = 1.0
    cfg.Uc = u_lb
    cfg.u_lb = u_lb
    cfg.tc = 1.0
    cfg.dt_d = 1.0
    cfg.force_factor = 1.0
    cfg.torque_factor = 1.0
    cfg.storage = N.STORAGE[storage] if isinstance(storage, str) else int(storage)
    cfg.device = device
    return cfg


class Lattice:
    """One LBM-PSM lattice resident on the GPU."""

    def __init__(self, cfg: N.Config):
        self.cfg = cfg
        h = C.c_void_p()
        N.check(N.LIB.lbmpsm_create(C.byref(cfg), C.byref(h)))
        self._h = h
        info = N.Info()
        self._check(N.LIB.lbmpsm_get_info(h, C.byref(info)))
        self.info = info
        self.dim, self.q = info.dim, info.q
        self.shape = (info.nx, info.ny, info.nz)
        self.nodes = info.nx * info.ny * info.nz
        self._n = 0

    # -- lifetime -------------------------------------------------------------
    def close(self):
        if self._h:
            N.LIB.lbmpsm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc):
        N.check(rc, self._h)

    # -- set-up ---------------------------------------------------------------
    def init_equilibrium(self, rho=1.0, u=(0.0, 0.0, 0.0)):
        self._check(N.LIB.lbmpsm_init_equilibrium(self._h, rho, u[0], u[1], u[2]))

    def set_bc(self, keyword, u_infty=None, rho_outlet=1.0, procgrid=None):
        """`fix lbm-psm-bc keyword`; the speed defaults to the fix's u_lb
        (src/fix_LBM_PSM_BC.cpp:156).  procgrid: the face ranges of a
        reference run on that rank grid (lbmpsm_set_bc_ranks)."""
        if u_infty is None:
            u_infty = self.cfg.u_lb
        kw = N.BC_KEYWORDS[keyword] if isinstance(keyword, str) else int(keyword)
        if procgrid is None:
            self._check(N.LIB.lbmpsm_set_bc(self._h, kw, u_infty, rho_outlet))
        else:
            pg = np.ascontiguousarray(procgrid, dtype=np.int32)
            self._check(N.LIB.lbmpsm_set_bc_ranks(self._h, kw, u_infty, rho_outlet, N.iptr(pg)))

    # -- one LBM step ---------------------------------------------------------
    def set_particles(self, tag, x, v, omega, radius):
        tag = np.ascontiguousarray(tag, dtype=np.int32)
        n = tag.shape[0]
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(n, 3)
        v = np.ascontiguousarray(v, dtype=np.float64).reshape(n, 3)
        w = np.ascontiguousarray(omega, dtype=np.float64).reshape(n, 3)
        r = np.ascontiguousarray(radius, dtype=np.float64).reshape(n)
        self._check(N.LIB.lbmpsm_set_particles(self._h, n, N.iptr(tag), N.dptr(x), N.dptr(v),
                                               N.dptr(w), N.dptr(r)))
        self._n = n

    def set_particles_device(self, n, tag_ptr, x_ptr, v_ptr, omega_ptr, radius_ptr, max_radius):
        """set_particles for a list already in device memory (raw device
        addresses of int32 tag[n] and float64 x/v/omega[n][3], radius[n])."""
        self._check(N.LIB.lbmpsm_set_particles_device(self._h, int(n), tag_ptr, x_ptr, v_ptr,
                                                      omega_ptr, radius_ptr, float(max_radius)))
        self._n = int(n)

    def step(self, store_macros=False):
        self._check(N.LIB.lbmpsm_step(self._h, N.STEP_STORE_MACROS if store_macros else 0))

    def hydro(self):
        n = self._n
        F = np.zeros((n, 3))
        T = np.zeros((n, 3))
        S = np.zeros((n, 6))
        self._check(N.LIB.lbmpsm_get_hydro(self._h, n, N.dptr(F), N.dptr(T), N.dptr(S)))
        return F, T, S

    # -- adjoint (frozen particles, FP64, fully periodic, no faces) ------------
    def adjoint_forward(self, steps, checkpoint_every, tag, x, v, omega, radius):
        """`steps` LBM steps with the same particle list mapped before each;
        returns the last step's per-entry F, T, S and keeps checkpoints for
        adjoint_backward."""
        tag = np.ascontiguousarray(tag, dtype=np.int32)
        n = tag.shape[0]
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(n, 3)
        v = np.ascontiguousarray(v, dtype=np.float64).reshape(n, 3)
        w = np.ascontiguousarray(omega, dtype=np.float64).reshape(n, 3)
        r = np.ascontiguousarray(radius, dtype=np.float64).reshape(n)
        F = np.zeros((n, 3))
        T = np.zeros((n, 3))
        S = np.zeros((n, 6))
        self._check(N.LIB.lbmpsm_adjoint_forward(
            self._h, int(steps), int(checkpoint_every), n, N.iptr(tag), N.dptr(x), N.dptr(v),
            N.dptr(w), N.dptr(r), N.dptr(F), N.dptr(T), N.dptr(S)))
        self._n = n
        return F, T, S

    def adjoint_backward(self, Fbar=None, Tbar=None, Sbar=None):
        """Gradients of a loss whose derivatives with respect to the last
        step's F, T, S are given: (d/dF_lbm [3], d/dtau, d/dv [n,3],
        d/domega [n,3])."""
        n = self._n

        def arr(a, cols):
            return None if a is None else np.ascontiguousarray(a, dtype=np.float64).reshape(n, cols)
        Fb, Tb, Sb = arr(Fbar, 3), arr(Tbar, 3), arr(Sbar, 6)
        gF = np.zeros(3)
        gtau = C.c_double()
        gv = np.zeros((n, 3))
        gw = np.zeros((n, 3))
        self._check(N.LIB.lbmpsm_adjoint_backward(
            self._h, n, None if Fb is None else N.dptr(Fb), None if Tb is None else N.dptr(Tb),
            None if Sb is None else N.dptr(Sb), N.dptr(gF), C.byref(gtau), N.dptr(gv),
            N.dptr(gw)))
        return gF, gtau.value, gv, gw

    # -- read-back ------------------------------------------------------------
    def populations(self):
        nx, ny, nz = self.shape
        f = np.empty((nx, ny, nz, self.q))
        self._check(N.LIB.lbmpsm_get_populations(self._h, N.dptr(f)))
        return f

    def set_populations(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64)
        assert f.size == self.nodes * self.q
        self._check(N.LIB.lbmpsm_set_populations(self._h, N.dptr(f)))

    def ref_shape(self):
        nx, ny, nz = self.shape
        return (nx + 2, ny + 2, nz + 2 if self.cfg.dim == 3 else 1, self.q)

    def populations_ref(self):
        """Populations in the reference's halo'd layout (restart files)."""
        f = np.empty(self.ref_shape())
        self._check(N.LIB.lbmpsm_get_populations_ref(self._h, N.dptr(f)))
        return f

    def set_populations_ref(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64)
        assert f.shape == self.ref_shape()
        self._check(N.LIB.lbmpsm_set_populations_ref(self._h, N.dptr(f)))

    def macros(self):
        nx, ny, nz = self.shape
        rho = np.empty((nx, ny, nz))
        u = np.empty((nx, ny, nz, 3))
        self._check(N.LIB.lbmpsm_get_macros(self._h, N.dptr(rho), N.dptr(u)))
        return rho, u

    def slots(self):
        nx, ny, nz = self.shape
        tag = np.empty((nx, ny, nz, 2), dtype=np.int32)
        sf = np.empty((nx, ny, nz, 2))
        us = np.empty((nx, ny, nz, 6))
        F = np.empty((nx, ny, nz, 6))
        self._check(N.LIB.lbmpsm_get_slots(self._h, N.iptr(tag), N.dptr(sf), N.dptr(us), N.dptr(F)))
        return tag, sf, us, F

    def mass(self):
        m = C.c_double()
        self._check(N.LIB.lbmpsm_mass(self._h, C.byref(m)))
        return m.value

    def sync(self):
        self._check(N.LIB.lbmpsm_sync(self._h))

    # -- timing ---------------------------------------------------------------
    def profile(self, enable=True):
        """True: time every kernel; "collide": the bulk collide only; False: off."""
        mode = 2 if enable == "collide" else (1 if enable else 0)
        self._check(N.LIB.lbmpsm_profile_enable(self._h, mode))

    def profile_reset(self):
        self._check(N.LIB.lbmpsm_profile_reset(self._h))

    def profile_read(self, kernel):
        ms = C.c_double()
        cnt = C.c_int64()
        self._check(N.LIB.lbmpsm_profile_read(self._h, kernel.encode(), C.byref(ms), C.byref(cnt)))
        return ms.value, cnt.value
