"""One rank against the reference's domain decomposition (CPU only).

The reference runs its Stokes drag deck (verification/3D_stokesDragForce-
tau0p7/in.lbmpsm: 100 x 101 x 100 nodes, `boundary p f p` with no
`lbm-psm-bc`, so the y halos stay as initialised) on `processors 2 2 2`.
This test runs that deck's first steps twice on the CPU oracle: on one
lattice, and on eight rank lattices that exchange their halos as
LBMPSMMPI::sendRecvData does (src/LBM_PSM_MPICOMM.h:79-147: per axis in
x, y, z order, whole halo'd planes, the last interior layer to the upper
neighbour and the first to the lower one, and no unpacking into the outer
halos of a non-periodic axis), with each rank's own local sizes
(src/LBM_PSM_lattice.cpp:20-52) and map / reduction offsets
(src/LBM_PSM_exchangeParticleData.cpp:25-35, 176-183), in the fix's order
(src/fix_LBM_PSM.cpp:196-213: map, halo exchange, collide-stream, then the
reduction of every atom on every rank, summed over ranks as
comm->reverse_comm() sums ghosts into owners).  The sphere starts at the box
centre, the corner all eight sub-domains share, and moves as
`fix move linear 0.02 0 0` moves it.

The decomposed run's interior populations equal the one-rank run's bit for
bit, and the drag agrees to summation order: the decomposition is not what
separates a one-rank run of this deck from the published 8-rank result.
"""
import itertools

import numpy as np

import lbmpsm_oracle as O
from lbmpsm import harness as H
from scenarios import check_hydro, hydro_scales

# verification/3D_stokesDragForce-tau0p7/in.lbmpsm:8-31
RHO, NU, LC, RE, NLC, TAU = 1000.0, 1e-4, 0.0005, 0.1, 11, 0.7
BOX = 0.005
N = (100, 101, 100)       # nodes: x, z periodic (100), y walls (101)
PER = (1, 0, 1)
GRID = (2, 2, 2)
STEPS = 30


def local_sizes(n, d):
    """Interior sizes of the ranks along one axis, src/LBM_PSM_lattice.cpp:20-52."""
    base = n // d
    return [base if (n % d == 0 or p != d - 1) else base + n - base * d for p in range(d)]


def take(axis, idx):
    s = [slice(None)] * 3
    s[axis] = idx
    return tuple(s)


def exchange(ranks, grid, periodic):
    """Halo exchange between rank lattices, src/LBM_PSM_MPICOMM.h:79-147."""
    views = {c: r.f_view() for c, r in ranks.items()}
    for a in range(3):
        d = grid[a]
        for c, f in views.items():  # lower halo <- lower neighbour's last interior layer
            if c[a] != 0 or periodic[a]:
                lo = list(c)
                lo[a] = (c[a] - 1) % d
                f[take(a, 0)] = views[tuple(lo)][take(a, -2)]
        for c, f in views.items():  # upper halo <- upper neighbour's first interior layer
            if c[a] != d - 1 or periodic[a]:
                hi = list(c)
                hi[a] = (c[a] + 1) % d
                f[take(a, -1)] = views[tuple(hi)][take(a, 1)]


def test_stokes_deck_two_by_two_by_two_ranks_equal_one_rank():
    unit = O.units(RHO, NU, LC, RE, NLC, TAU, 3)
    box = H.Box(lo=np.zeros(3), prd=np.array([BOX, BOX, BOX]), periodic=PER)
    atoms = H.Atoms(tag=np.array([1], np.int32), x=np.array([[0.0025, 0.0025, 0.0025]]),
                    v=np.array([[0.02, 0.0, 0.0]]), omega=np.zeros((1, 3)),
                    radius=np.array([0.00025]))
    dt = unit.dt_d * unit.tc

    one = O.Oracle(3, N, PER, TAU, unit=unit)
    sizes = [local_sizes(N[a], GRID[a]) for a in range(3)]
    ranks = {}
    for c in itertools.product(*[range(g) for g in GRID]):
        n_loc = [sizes[a][c[a]] for a in range(3)]
        r = O.Oracle(3, n_loc, (0, 0, 0), TAU, unit=unit)
        r.set_rank([sum(s + 2 for s in sizes[a][:c[a]]) for a in range(3)], c)
        ranks[c] = r
    starts = [np.concatenate(([0], np.cumsum(sizes[a])))[:-1] for a in range(3)]

    for step in range(1, STEPS + 1):
        H.move_linear(atoms, box, step, dt)
        lst = H.build_list(atoms, box, cutoff=0.001)  # comm_modify cutoff 0.001
        F1, T1, S1 = one.fix_step(lst)
        for r in ranks.values():
            r.map(lst)
        exchange(ranks, GRID, PER)
        Fr, Tr, Sr = np.zeros_like(F1), np.zeros_like(T1), np.zeros_like(S1)
        for r in ranks.values():
            r.collide_stream()
        for c in sorted(ranks):
            F, Tq, S = ranks[c].hydro(lst)
            Fr += F
            Tr += Tq
            Sr += S
        # the sums differ in order only: per component within 1e-12 relative,
        # or of the bound on the sum of |terms| / 100 where they cancel
        tag, _, _, Fnode = one.slots()
        sF, sT = hydro_scales(lst, tag, Fnode, unit.forceFactor, unit.torqueFactor, unit.dx)
        check_hydro(Fr, F1, sF, label=f"step {step} force")
        check_hydro(Tr, T1, sT, label=f"step {step} torque")
        check_hydro(Sr, S1, sT, label=f"step {step} stresslet")
        fx1, fxr = F1[0, 0], Fr[0, 0]

    f1 = one.populations()
    for c, r in ranks.items():
        sl = tuple(slice(starts[a][c[a]], starts[a][c[a]] + sizes[a][c[a]]) for a in range(3))
        assert np.array_equal(r.populations(), f1[sl]), c
    print(f"\n  {STEPS} steps: 2x2x2 ranks equal one rank bit for bit; drag {fx1:.15e} "
          f"(ranks {fxr:.15e})")
    one.close()
    for r in ranks.values():
        r.close()
