"""A short fluidised-bed column with `pair_style lubricate/GRM/LBDEM` active,
driven through the mock LAMMPS: `fix lbm-psm` + `fix lbm-psm-bc xFlow` (the
C++ fixes over the GPU library) and the lubrication pair (host C++), against
the CPU oracle of the lattice step plus oracle/grm_oracle.py of the pair.

Column 96 x 48 x 48 nodes (x vertical, `boundary f p p`), Nlc 11 with the
sphere diameter lc = 1 mm (dx = 0.1 mm, d = 10 nodes), water (rho 1000,
nu 1e-6), Re 0.5, tau 0.7, gravity Fext along -x; 40 spheres by RSA, among
them 12 pairs placed at gaps of 0.02 to 0.9 nodes, all moving linearly
(`fix move linear`, speeds up to 0.3 Uc).  Squeeze lubrication (flagGRM 1)
acts below a gap of one node (cut 0.1 mm x cutoff factor 1).  Per step:
the LBM force and torque on every sphere within 1e-12 relative per component
(the oracle's sum of term magnitudes / 100 for cancellation-dominated
components), the pair forces bit-identical, and their sum (what LAMMPS hands
to the integrator) within the same bar.  Reference:
src/fix_LBM_PSM.cpp:181-266, src/pair_GRM_LBDEM.cpp:78-476.
"""
import math
import os
import sys

import numpy as np
import pytest

import lbmpsm_oracle as O
from lammps_mock import Lammps
from lbmpsm import harness as H
from scenarios import check_hydro, hydro_scales, owner_scales

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                "oracle"))
import grm_oracle as G  # noqa: E402

pytestmark = pytest.mark.gpu

LC, NLC, RHO, NU, RE, TAU = 0.001, 11, 1000.0, 1e-6, 0.5, 0.7
DX = LC / (NLC - 1)
N = (96, 48, 48)
FEXT = -20.0  # N/m^3 along x


def _spheres(rng, box):
    r = LC / 2.0
    pos = []

    def free(c):
        for p in pos:
            d = c - p
            for a in (1, 2):
                d[a] -= box.prd[a] * round(d[a] / box.prd[a])
            if np.linalg.norm(d) < 2 * r + 0.005 * DX:
                return False
        return True

    lo, hi = 6 * DX + r, box.prd[0] - 6 * DX - r
    while len(pos) < 24:  # 12 close pairs
        c = np.array([rng.uniform(lo, hi), rng.uniform(0, box.prd[1]), rng.uniform(0, box.prd[2])])
        e = rng.normal(size=3)
        e /= np.linalg.norm(e)
        c2 = c + e * (2 * r + rng.uniform(0.02, 0.9) * DX)
        c2[1:] %= box.prd[1:]
        if lo < c2[0] < hi and free(c) and free(c2):
            pos += [c, c2]
    while len(pos) < 40:
        c = np.array([rng.uniform(lo, hi), rng.uniform(0, box.prd[1]), rng.uniform(0, box.prd[2])])
        if free(c):
            pos.append(c)
    return np.array(pos)


def _neigh(lst, reach):
    """Full neighbour list of the locals over the whole atom list."""
    out = []
    for i in range(lst.nlocal):
        d = np.linalg.norm(lst.x - lst.x[i], axis=1)
        out.append([int(j) for j in np.argsort(d, kind="stable") if j != i and d[j] < reach])
    return out


def test_bed_column_with_lubrication_pair():
    box = H.Box(lo=np.zeros(3), prd=np.array([(N[0] - 1) * DX, N[1] * DX, N[2] * DX]),
                periodic=(0, 1, 1), dim=3)
    rng = np.random.default_rng(2024)
    pos = _spheres(rng, box)
    n = len(pos)
    Uc = RE * NU / LC
    atoms = H.Atoms(tag=np.arange(1, n + 1, dtype=np.int32), x=pos,
                    v=rng.uniform(-0.3, 0.3, (n, 3)) * Uc,
                    omega=rng.uniform(-0.3, 0.3, (n, 3)) * Uc / LC,
                    radius=np.full(n, LC / 2.0))
    lmp = Lammps(box)
    lmp.fix(f"1 all lbm-psm every 1 Nlc {NLC} lc {LC} rho {RHO} nu {NU} Re {RE} tau {TAU} "
            f"Fext {FEXT} 0 0")
    lmp.fix("2 all lbm-psm-bc xFlow")
    lmp.pair_env(ntypes=1, vxmu2f=1.0)
    mu = RHO * NU
    lmp.pair_style(f"lubricate/GRM/LBDEM {mu} {0.001 * DX} {DX} 1 1.0")
    lmp.pair_coeff("1 1")
    lmp.init()
    dt = lmp.dt
    u = O.units(RHO, NU, LC, RE, NLC, TAU, 3)
    Flb = FEXT / u.volumeForceFactor
    ora = O.Oracle(3, N, box.periodic, TAU, F_lbm=(Flb, 0.0, 0.0), unit=u, origin=box.lo)
    print(f"\n[case] bed column {N} xFlow, {n} spheres (12 close pairs), GRM pair, 8 steps",
          flush=True)
    close_seen = 0
    for step in range(1, 9):
        H.move_linear(atoms, box, step, dt)
        lst = H.build_list(atoms, box, cutoff=LC)
        f_lb, t_lb, _ = lmp.step(lst)
        neigh = _neigh(lst, 2 * LC)
        f_p, t_p, _ = lmp.pair_compute(lst.x, lst.v, lst.omega, lst.radius,
                                       np.ones(lst.n, np.int32), lst.nlocal, neigh)
        F, T, S = ora.fix_step(lst, bc_keyword=2)
        Fo, To = H.reduce_to_owners(lst, F, T)
        to, _, _, Fho = ora.slots()
        oF, oT = owner_scales(lst, *hydro_scales(lst, to, Fho, u.forceFactor, u.torqueFactor,
                                                 u.dx))
        check_hydro(f_lb, Fo, oF, 1e-12, f"step {step} LBM force")
        check_hydro(t_lb, To, oT, 1e-12, f"step {step} LBM torque")
        L = lambda a: np.asarray(a).tolist()  # noqa: E731
        fpo, tpo, _ = G.compute(L(lst.x), L(lst.v), L(lst.omega), L(lst.radius),
                                [1] * lst.n, lst.nlocal, neigh, mu=mu, flag=1,
                                cut={(1, 1): DX}, cut_inner={(1, 1): 0.001 * DX},
                                cutoff_factor=1.0, vxmu2f=1.0)
        fpo = np.array(fpo)
        assert np.array_equal(f_p, fpo), step
        close_seen = max(close_seen, int(np.count_nonzero(np.linalg.norm(fpo, axis=1))))
        check_hydro(f_lb + f_p, Fo + fpo, oF + np.linalg.norm(fpo, axis=1), 1e-12,
                    f"step {step} total force")
    assert close_seen >= 12  # the lubrication term acts on the close pairs
    assert np.max(np.abs(f_lb)) > 0.0
    ora.close()
    lmp.close()
