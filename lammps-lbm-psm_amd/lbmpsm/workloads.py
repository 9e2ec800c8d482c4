"""Synthetic particle workloads in lattice units (dx = 1), seeded.

Each builder returns (box, atoms) for the mock-LAMMPS harness: a LAMMPS box
with the node counts of `fix lbm-psm` (a periodic axis of n nodes has length
n, a wall axis n - 1, src/fix_LBM_PSM.cpp:142-154) and the locals of the
atom list; `atom_lists` adds prescribed linear motion and periodic ghost
images per step.

* shear_suspension: D3Q19 192^3 with shear walls in y (`boundary p f p`),
  400 spheres d = 16 by random sequential addition (RSA), v ~ U(+-1e-3),
  omega ~ U(+-1e-4).
* fluidised_bed: D3Q19 1024 x 256 x 256 column, x vertical (`boundary f p p`,
  `lbm-psm-bc xFlow`), 20000 spheres d = 12 by RSA over the column away from
  the inlet and outlet (phi = 0.27; 20000 spheres in its lower third would
  need phi = 0.81, beyond RSA's jamming limit of ~0.38), v ~ U(+-1e-3),
  omega ~ U(+-1e-4).
* dense_packing: D3Q19 256^3 periodic, d ~ U[10, 20] to phi = 0.30, largest
  first, fixed.
"""
from __future__ import annotations

import numpy as np

from . import harness as H


def rsa_cells(L, r, n, seed, lo_x, hi_x, periodic_x=False):
    """Random sequential addition of n equal spheres of radius r in a box of
    lengths L, periodic in y and z (and x if periodic_x), centres in
    [lo_x, hi_x) along x; a cell grid keeps the overlap test local."""
    rng = np.random.default_rng(seed)
    cell = 2.0 * r
    d2min = (2.0 * r) ** 2
    Lx, Ly, Lz = (float(v) for v in L[:3])
    nc = [max(1, int(L[d] // cell)) for d in range(3)]
    per = (bool(periodic_x), True, True)
    grid = {}
    pos = []
    tries = 0
    while len(pos) < n:
        tries += 1
        if tries > 50 * n:
            raise RuntimeError(f"RSA jammed after placing {len(pos)} of {n} spheres")
        cx = rng.uniform(lo_x, hi_x)
        cy = rng.uniform(0, Ly)
        cz = rng.uniform(0, Lz)
        ci = (int(cx // cell), int(cy // cell), int(cz // cell))
        ok = True
        for ox in (-1, 0, 1):
            kx = (ci[0] + ox) % nc[0] if per[0] else ci[0] + ox
            for oy in (-1, 0, 1):
                ky = (ci[1] + oy) % nc[1]
                for oz in (-1, 0, 1):
                    for (px, py, pz) in grid.get((kx, ky, (ci[2] + oz) % nc[2]), ()):
                        dx = cx - px
                        if per[0]:
                            dx -= Lx * round(dx / Lx)
                        dy = cy - py
                        dy -= Ly * round(dy / Ly)
                        dz = cz - pz
                        dz -= Lz * round(dz / Lz)
                        if dx * dx + dy * dy + dz * dz < d2min:
                            ok = False
                            break
                    if not ok:
                        break
                if not ok:
                    break
            if not ok:
                break
        if ok:
            key = (ci[0] % nc[0] if per[0] else ci[0], ci[1] % nc[1], ci[2] % nc[2])
            grid.setdefault(key, []).append((cx, cy, cz))
            pos.append((cx, cy, cz))
    return np.array(pos)


def _moving(pos, radius, seed, vmax=1e-3, wmax=1e-4):
    n = pos.shape[0]
    rng = np.random.default_rng(seed + 1)
    v = rng.uniform(-vmax, vmax, (n, 3))
    w = rng.uniform(-wmax, wmax, (n, 3))
    return H.Atoms(tag=np.arange(1, n + 1, dtype=np.int32), x=pos, v=v, omega=w,
                   radius=np.full(n, float(radius)) if np.ndim(radius) == 0 else radius)


SHEAR = dict(n=(192, 192, 192), periodic=(1, 0, 1), tau=0.7, u_wall=0.01, spheres=400,
             radius=8.0, seed=20240611, bc="shear")


def shear_suspension(spheres=SHEAR["spheres"], seed=SHEAR["seed"]):
    box = H.Box(lo=np.zeros(3), prd=np.array([192.0, 191.0, 192.0]), periodic=(1, 0, 1), dim=3)
    radii = np.full(spheres, SHEAR["radius"])
    pos = H.rsa_spheres(box, radii, seed, gap=2.5)
    return box, _moving(pos, SHEAR["radius"], seed)


BED = dict(n=(1024, 256, 256), periodic=(0, 1, 1), tau=0.7, u_inlet=0.005, F=(-2e-6, 0.0, 0.0),
           spheres=20000, radius=6.0, seed=505, bc="xFlow")


def fluidised_bed(spheres=BED["spheres"], seed=BED["seed"], n=BED["n"]):
    box = H.Box(lo=np.zeros(3), prd=np.array([n[0] - 1.0, float(n[1]), float(n[2])]),
                periodic=BED["periodic"], dim=3)
    pos = rsa_cells(box.prd, BED["radius"], spheres, seed, 20.0, box.prd[0] - 20.0)
    return box, _moving(pos, BED["radius"], seed)


def dense_packing(seed=303, n=256, phi=0.30):
    box = H.Box(lo=np.zeros(3), prd=np.array([float(n)] * 3), periodic=(1, 1, 1), dim=3)
    rng = np.random.default_rng(seed)
    d = rng.uniform(10.0, 20.0, 4000)
    vol = np.pi / 6.0 * np.cumsum(d ** 3)
    m = int(np.searchsorted(vol, phi * n ** 3)) + 1
    d = np.sort(d[:m])[::-1]
    pos = H.rsa_spheres(box, d / 2.0, seed, gap=0.0)
    atoms = H.Atoms(tag=np.arange(1, m + 1, dtype=np.int32), x=pos, v=np.zeros((m, 3)),
                    omega=np.zeros((m, 3)), radius=d / 2.0)
    return box, atoms


def atom_lists(box, atoms, nsteps, cutoff, first=0):
    """Per-step atom lists (locals + ghost images) under `fix move linear`
    with dt = 1, steps first .. first + nsteps - 1."""
    a = atoms.copy()
    out = []
    for s in range(first, first + nsteps):
        H.move_linear(a, box, s, 1.0)
        out.append(H.build_list(a, box, cutoff=cutoff))
    return out
