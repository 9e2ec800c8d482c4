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


def hydro_scales(lst, slot_tag, slot_F, force_factor=1.0, torque_factor=1.0):
    """Upper bound of the sum of |terms| in each entry's force and torque
    reductions: all node forces of the entry's tag (owner and images share
    it), times the largest lever arm for torque and stresslet.  Reordering a
    sum changes it by at most ~n*eps times this bound, so a tolerance of
    1e-12 of it is order-independent."""
    mag = np.linalg.norm(slot_F.reshape(-1, 2, 3), axis=2)  # [node, slot]
    tags = slot_tag.reshape(-1, 2)
    sF = np.zeros(lst.n)
    for p in range(lst.n):
        t = lst.tag[p]
        sF[p] = np.sum(mag[tags[:, 0] == t, 0]) + np.sum(mag[tags[:, 1] == t, 1])
    sF = sF * force_factor
    sT = sF / force_factor * torque_factor * (lst.radius + 4.0)
    return sF, sT


def rel_err(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    den = np.max(np.abs(b))
    if den == 0:
        return float(np.max(np.abs(a)))
    return float(np.max(np.abs(a - b)) / den)
