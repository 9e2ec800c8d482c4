"""Shared scenario builders for the parity tests: the same mock-LAMMPS atom
lists are fed to the GPU library and to the CPU oracle."""
from __future__ import annotations

import numpy as np

from lbmpsm import harness as H

BC_IDS = {"none": 0, "shear": 1, "xFlow": 2, "channel": 3, "channel-velIn-pressOut": 4,
          "closedBox": 5}


def lattice_box(dim, n, periodic):
    """Lattice-unit box: dx = 1, node i at x = i; a wall axis of n nodes has
    length n - 1 (src/fix_LBM_PSM.cpp:142-154)."""
    prd = np.array([float(n[d]) if periodic[d] else float(n[d] - 1) for d in range(3)])
    if dim == 2:
        prd[2] = 1.0
    return H.Box(lo=np.zeros(3), prd=prd, periodic=tuple(periodic), dim=dim)


def random_atoms(box, radii, seed, vmax=1e-3, wmax=1e-4, gap=0.3, positions=None):
    rng = np.random.default_rng(seed + 1000)
    n = len(radii)
    pos = positions if positions is not None else H.rsa_spheres(box, np.asarray(radii, float),
                                                                 seed, gap=gap)
    v = rng.uniform(-vmax, vmax, (n, 3))
    w = rng.uniform(-wmax, wmax, (n, 3))
    if box.dim == 2:
        pos[:, 2] = 0.0
        v[:, 2] = 0.0
        w[:, 0] = 0.0
        w[:, 1] = 0.0
    return H.Atoms(tag=np.arange(1, n + 1, dtype=np.int32), x=pos.copy(), v=v, omega=w,
                   radius=np.asarray(radii, float))


def hydro_scales(lst, slot_tag, slot_F, force_factor=1.0, torque_factor=1.0, dx=1.0):
    """Upper bound of the sum of |terms| in each entry's force and torque
    reductions: all node forces of the entry's tag (owner and images share
    it), times the largest lever arm for torque and stresslet.  Reordering a
    sum changes it by at most ~n*eps times this bound, so a tolerance of
    1e-12 of it is order-independent."""
    mag = np.linalg.norm(np.asarray(slot_F).reshape(-1, 2, 3), axis=2)  # [node, slot]
    tags = np.asarray(slot_tag).reshape(-1, 2)
    top = int(max(np.max(lst.tag, initial=0), np.max(tags, initial=0))) + 1
    per_tag = np.zeros(top)
    for s in range(2):
        held = tags[:, s] > 0
        per_tag += np.bincount(tags[held, s], weights=mag[held, s], minlength=top)
    sF = per_tag[np.asarray(lst.tag)] * force_factor
    sT = sF / force_factor * torque_factor * (lst.radius / dx + 4.0)
    return sF, sT


def owner_scales(lst, sF, sT):
    """Per-owner bounds after the reverse communication: the largest bound
    among an owner's entries (its images share its tag and node forces)."""
    oF = np.zeros(lst.nlocal)
    oT = np.zeros(lst.nlocal)
    np.maximum.at(oF, lst.owner, sF)
    np.maximum.at(oT, lst.owner, sT)
    return oF, oT


def rel_err(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    den = np.max(np.abs(b))
    if den == 0:
        return float(np.max(np.abs(a)))
    return float(np.max(np.abs(a - b)) / den)


# A force, torque or stresslet component is "cancellation-dominated" when its
# value is more than K_CANCEL times smaller than the bound on the sum of the
# absolute values of its terms (or, without that bound, than the entry's
# largest component of the same quantity).  Reordering a sum of n terms
# moves it by ~n eps times that bound, so only for the others is a relative
# bar order-independent.
K_CANCEL = 100.0


def check_hydro(got, ref, scale=None, tol=1e-12, label=""):
    """Per entry and component, |got - ref| <= tol |ref| where the value is
    not cancellation-dominated, and <= tol scale / K_CANCEL where it is.
    got/ref: [n, c]; scale: [n] bound of the sum of |terms| per entry, or
    None for the entry's largest |ref| component.  Prints and returns the
    largest relative error among the non-cancelling components and the count
    of cancelling ones."""
    got = np.asarray(got, float)
    ref = np.asarray(ref, float)
    if got.size == 0:
        return 0.0, 0
    if scale is None:
        scale = np.max(np.abs(ref), axis=1)
    floor = np.asarray(scale, float)[:, None] / K_CANCEL
    d = np.abs(got - ref)
    mag = np.abs(ref)
    big = mag >= floor
    rel = np.where(big, d / np.where(mag > 0, mag, 1.0), 0.0)
    worst = float(rel.max()) if big.any() else 0.0
    ncancel = int((~big).sum())
    print(f"  {label} max rel err {worst:.2e} ({big.sum()} comps), "
          f"{ncancel} cancellation-dominated comps within {tol:.0e} x scale/{K_CANCEL:.0f}",
          flush=True)
    assert np.all(d <= tol * np.maximum(mag, floor) + 1e-300), label
    return worst, ncancel


# Velocity sets in the reference's population order
# (src/LBM_PSM_lattice.h:186-209) and the second-order equilibrium, for
# building non-uniform initial states in tests.
E9 = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [-1, 0, 0], [0, -1, 0],
               [1, 1, 0], [-1, 1, 0], [-1, -1, 0], [1, -1, 0]], float)
W9 = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)
E19 = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
                [1, 1, 0], [-1, -1, 0], [1, 0, 1], [-1, 0, -1], [0, 1, 1], [0, -1, -1],
                [1, -1, 0], [-1, 1, 0], [1, 0, -1], [-1, 0, 1], [0, 1, -1], [0, -1, 1]], float)
W19 = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)


def equilibrium(dim, rho, u):
    """f_eq = w rho (1 + 3 e.u + 4.5 (e.u)^2 - 1.5 u^2) per node; rho (...),
    u (..., 3) -> (..., q)."""
    E, W = (E19, W19) if dim == 3 else (E9, W9)
    eu = u @ E.T
    usq = np.sum(u * u, axis=-1)[..., None]
    return W * rho[..., None] * (1.0 + 3.0 * eu + 4.5 * eu * eu - 1.5 * usq)
