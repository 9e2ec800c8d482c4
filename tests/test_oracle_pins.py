"""Pins of the CPU oracle (no GPU needed).

The reference ships no unit-level golden vectors, so the oracle is checked
against what the mathematics fixes, chosen so that a plausible slip (a wrong
weight, sign, index or prefactor) fails at least one test:

* lattice isotropy of the reference velocity sets and the moments of the
  equilibrium built from them;
* exact invariants: rest state, mass conservation, Guo momentum gain, the
  fluid-particle momentum balance;
* closed-form steady flows: Couette (shear walls) and Poiseuille (channel
  + body force);
* Zou-He nodes reproduce the imposed velocity / density;
* sub-sampled solid fraction vs sphere volume and a brute-force count;
* hand-built slot state-machine scenarios and known-answer torque /
  stresslet sums.
"""
import ctypes as C
import math

import numpy as np
import pytest

import lbmpsm_oracle as O
from lbmpsm import harness as H
from scenarios import lattice_box, random_atoms

# Reference velocity sets, src/LBM_PSM_lattice.h:186-209.
D2Q9 = dict(
    e=np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [-1, 0, 0], [0, -1, 0], [1, 1, 0], [-1, 1, 0],
                [-1, -1, 0], [1, -1, 0]]),
    w=np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4),
)
D3Q19 = dict(
    e=np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
                [1, 1, 0], [-1, -1, 0], [1, 0, 1], [-1, 0, -1], [0, 1, 1], [0, -1, -1],
                [1, -1, 0], [-1, 1, 0], [1, 0, -1], [-1, 0, 1], [0, 1, -1], [0, -1, 1]]),
    w=np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12),
)

EMPTY = H.AtomList(tag=np.zeros(0, np.int32), x=np.zeros((0, 3)), v=np.zeros((0, 3)),
                   omega=np.zeros((0, 3)), radius=np.zeros(0), owner=np.zeros(0, int), nlocal=0)


def moments(f, lat):
    e = lat["e"].astype(float)
    rho = f.sum(axis=-1)
    j = f @ e
    P = np.einsum("...q,qa,qb->...ab", f, e, e)
    return rho, j, P


@pytest.mark.parametrize("lat", [D2Q9, D3Q19], ids=["D2Q9", "D3Q19"])
def test_velocity_set_isotropy(lat):
    e, w = lat["e"].astype(float), lat["w"]
    dim = 2 if len(w) == 9 else 3
    assert abs(w.sum() - 1) < 1e-15
    assert np.allclose(w @ e, 0)
    assert np.allclose(np.einsum("q,qa,qb->ab", w, e, e)[:dim, :dim], np.eye(dim) / 3)
    m4 = np.einsum("q,qa,qb,qc,qd->abcd", w, e, e, e, e)
    d = np.eye(3)
    iso = (np.einsum("ab,cd->abcd", d, d) + np.einsum("ac,bd->abcd", d, d)
           + np.einsum("ad,bc->abcd", d, d)) / 9
    assert np.allclose(m4[:dim, :dim, :dim, :dim], iso[:dim, :dim, :dim, :dim])
    # opposite directions used by Zou-He: i+1 / i-1 in D3Q19, fixed pairs in D2Q9
    q = len(w)
    for i in range(1, q):
        opp = [k for k in range(q) if np.all(lat["e"][k] == -lat["e"][i])]
        assert len(opp) == 1 and w[opp[0]] == w[i]
        if q == 19:
            assert opp[0] == (i + 1 if i % 2 == 1 else i - 1)


@pytest.mark.parametrize("dim,lat", [(2, D2Q9), (3, D3Q19)], ids=["D2Q9", "D3Q19"])
def test_equilibrium_moments(dim, lat):
    n = (6, 5, 4)
    o = O.Oracle(dim, n, (1, 1, 1), 0.8)
    rho, u = 1.03, np.array([0.02, -0.013, 0.007 if dim == 3 else 0.0])
    o.initialise(rho, u)
    f = o.populations()
    r, j, P = moments(f, lat)
    assert np.allclose(r, rho, rtol=1e-15, atol=0)
    assert np.allclose(j, rho * u, rtol=0, atol=1e-16)
    expect = rho * (np.eye(3) / 3 + np.outer(u, u))
    expect[2, 2] = 0.0 if dim == 2 else expect[2, 2]
    assert np.allclose(P[0, 0, 0], expect, atol=1e-15)


@pytest.mark.parametrize("bc,per", [("none", (1, 1, 1)), ("channel", (1, 0, 1)),
                                    ("closedBox", (0, 0, 0))])
def test_rest_state_is_invariant(bc, per):
    """A fluid at rest stays at rest: sum w_i rounds to 1 +- 1 ulp, so the
    populations may move by a few ulps but never drift."""
    ids = {"none": 0, "channel": 3, "closedBox": 5}
    o = O.Oracle(3, (8, 7, 6), per, 0.9)
    f0 = o.populations()
    for _ in range(50):
        o.fix_step(EMPTY, bc_keyword=ids[bc])
    assert np.max(np.abs(o.populations() - f0) / f0) < 1e-14
    _, u = o.macros()
    assert np.max(np.abs(u)) < 1e-15


def test_mass_conserved_periodic_with_moving_particles():
    n = (20, 18, 16)
    box = lattice_box(3, n, (1, 1, 1))
    atoms = random_atoms(box, [3.5, 3.0, 4.0], seed=4, vmax=0.02, wmax=5e-3, gap=0.1)
    o = O.Oracle(3, n, (1, 1, 1), 0.75)
    rng = np.random.default_rng(0)
    fh = o.f_halo()
    fh *= 1 + 0.01 * rng.standard_normal(fh.shape)
    o.set_f_halo(fh)
    m0 = math.fsum(o.populations().ravel())
    for s in range(15):
        H.move_linear(atoms, box, s, 1.0)
        o.fix_step(H.build_list(atoms, box, 8.0))
    m1 = math.fsum(o.populations().ravel())
    assert abs(m1 - m0) <= 1e-12 * m0


@pytest.mark.parametrize("dim", [2, 3])
def test_guo_momentum_gain_is_F_per_step(dim):
    n = (6, 5, 4)
    F = np.array([1e-5, -3e-6, 2e-6 if dim == 3 else 0.0])
    o = O.Oracle(dim, n, (1, 1, 1), 0.7, F_lbm=F)
    nodes = n[0] * n[1] * (n[2] if dim == 3 else 1)
    lat = D3Q19 if dim == 3 else D2Q9
    steps = 25
    for _ in range(steps):
        o.fix_step(EMPTY)
    _, j, _ = moments(o.populations(), lat)
    # |j| ~ 1e-2 summed from ~2000 populations of ~5e-2: rounding ~1e-14
    assert np.allclose(j.reshape(-1, 3).sum(0), steps * F * nodes, rtol=0, atol=1e-13)
    # stored velocity carries the half-force shift: u = (j + F/2)/rho
    _, u = o.macros()
    assert np.allclose(u.reshape(-1, 3), (steps - 1 + 0.5) * F, rtol=1e-10, atol=1e-18)


def test_momentum_balance_fluid_particles():
    """Fluid momentum change in one step = - sum of node forces on all slots
    (src/LBM_PSM_BGK_dynamics.h:98-113 define both from the same Omega)."""
    n = (24, 24, 24)
    box = lattice_box(3, n, (1, 1, 1))
    pos = np.array([[9.3, 11.8, 12.4], [15.1, 12.6, 11.2]])
    atoms = random_atoms(box, [3.2, 2.9], seed=2, positions=pos, vmax=0.01, wmax=2e-3)
    o = O.Oracle(3, n, (1, 1, 1), 0.8)
    o.initialise(1.0, (0.01, -0.004, 0.002))
    lst = H.build_list(atoms, box, 3.5)  # zones stay away from the periodic seams
    assert lst.n == 2
    o.fix_step(lst)  # warm the slots, then measure one step
    _, j0, _ = moments(o.populations(), D3Q19)
    F, T, S = o.fix_step(lst)
    _, j1, _ = moments(o.populations(), D3Q19)
    _, _, _, Fh = o.slots()
    dj = j1.reshape(-1, 3).sum(0) - j0.reshape(-1, 3).sum(0)
    Fnodes = Fh.reshape(-1, 6)[:, :3].sum(0) + Fh.reshape(-1, 6)[:, 3:].sum(0)
    assert np.allclose(dj, -Fnodes, rtol=1e-9, atol=1e-15)
    assert np.allclose(F.sum(0), Fnodes, rtol=1e-9, atol=1e-15)
    assert np.linalg.norm(Fnodes) > 1e-6


def test_couette_profile_linear():
    n = (4, 21, 1)
    o = O.Oracle(2, n, (1, 0, 1), 0.8, unit=O.lattice_units(0.01))
    for _ in range(6000):
        o.fix_step(EMPTY, bc_keyword=1)
    _, u = o.macros()
    ux = u[0, :, 0, 0]
    y = np.arange(21)
    expect = -0.01 + 0.02 * y / 20.0
    assert np.max(np.abs(ux - expect)) < 1e-7


def test_poiseuille_profile():
    n = (4, 19, 1)
    F = 1e-6
    tau = 0.8
    nu = (tau - 0.5) / 3
    o = O.Oracle(2, n, (1, 0, 1), tau, F_lbm=(F, 0, 0), unit=O.lattice_units(0.01))
    for _ in range(8000):
        o.fix_step(EMPTY, bc_keyword=3)
    _, u = o.macros()
    ux = u[0, :, 0, 0]
    y = np.arange(19)
    expect = F / (2 * nu) * y * (18 - y)
    assert np.max(np.abs(ux - expect)) < 2e-3 * np.max(expect)


def test_zou_he_imposes_velocity_and_density():
    n = (10, 8, 7)
    U = 0.01
    o = O.Oracle(3, n, (0, 1, 1), 0.7, unit=O.lattice_units(U))
    rng = np.random.default_rng(1)
    fh = o.f_halo()
    fh *= 1 + 0.02 * rng.standard_normal(fh.shape)
    o.set_f_halo(fh)
    o.fix_step(EMPTY, bc_keyword=2)  # xFlow: xn velocity U, xp density 1
    f = o.populations()
    rho, j, _ = moments(f, D3Q19)
    # interior of the inlet face (away from the y/z faces that follow it)
    inlet = (j[1, 1:-1, 1:-1, 0] / rho[1, 1:-1, 1:-1])
    # the x = 0 global node is the first interior node of the reference box
    ux_in = j[0, 1:-1, 1:-1, 0] / rho[0, 1:-1, 1:-1]
    assert np.allclose(ux_in, U, rtol=1e-12, atol=1e-15)
    assert np.allclose(rho[-1, 1:-1, 1:-1], 1.0, rtol=1e-14)
    assert inlet.size  # the node behind the inlet is not constrained


def test_solid_fraction_volume_and_brute_force():
    o = O.Oracle(3, (30, 30, 30), (1, 1, 1), 0.8)
    rng = np.random.default_rng(3)
    hs = math.sqrt(3.1) / 2.0
    for r in (3.0, 5.5, 8.0):
        c = np.array([15.0, 15.0, 15.0]) + rng.uniform(-0.5, 0.5, 3)
        xl = c + 1.0  # reference halo'd coordinates
        tot = 0.0
        for i in range(int(xl[0] - r) - 2, int(xl[0] + r) + 3):
            for j in range(int(xl[1] - r) - 2, int(xl[1] + r) + 3):
                for k in range(int(xl[2] - r) - 2, int(xl[2] + r) + 3):
                    tot += o.solid_fraction(i, j, k, xl[0], xl[1], xl[2], r)
        vol = 4.0 / 3.0 * math.pi * r ** 3
        assert abs(tot - vol) < 0.01 * vol + 0.5
    # brute-force sub-sampling on random rim cells
    delta = -0.5 + (np.arange(10) + 0.5) * (1.0 / 10.0)
    for _ in range(300):
        r = rng.uniform(1.0, 10.0)
        d = rng.standard_normal(3)
        d = d / np.linalg.norm(d) * (r + rng.uniform(-hs, hs))
        xl = np.array([12.0, 12.0, 12.0]) - d
        got = o.solid_fraction(12, 12, 12, xl[0], xl[1], xl[2], r)
        dd = np.array([12.0, 12.0, 12.0]) - xl
        dist = float(dd @ dd)
        if dist > (r + hs) ** 2:
            expect = 0.0
        elif dist < (r - hs) ** 2:
            expect = 1.0
        else:
            X = (dd[0] + delta) ** 2
            Y = (dd[1] + delta) ** 2
            Z = (dd[2] + delta) ** 2
            cnt = int(np.sum((X[:, None, None] + Y[None, :, None]) + Z[None, None, :] < r * r))
            expect = (1.0 / 1000.0) * cnt
        assert got == expect


def _node(o, i, j, k):
    """Reference halo'd index -> interior-array index of the oracle slots."""
    return (i - 1, j - 1, k - 1)


def test_slot_state_machine_scenarios():
    """Three spheres whose rims meet at one node: the first two take slots 0
    and 1 in list order, the third is dropped there; when the first leaves,
    its slot is emptied and the next newcomer takes it
    (src/LBM_PSM_lattice.cpp:257-324, src/LBM_PSM_exchangeParticleData.cpp:81-106)."""
    n = (20, 20, 20)
    box = lattice_box(3, n, (1, 1, 1))
    o = O.Oracle(3, n, (1, 1, 1), 0.8)
    node = np.array([10.0, 10.0, 10.0])
    r = 2.5
    dirs = np.array([[1, 0, 0], [-0.5, 0.866, 0], [-0.5, -0.866, 0]])
    pos = node + (r + 0.2) * dirs  # rims straddle the node
    atoms = H.Atoms(tag=np.array([7, 3, 5], np.int32), x=pos, v=np.zeros((3, 3)),
                    omega=np.zeros((3, 3)), radius=np.full(3, r))
    o.fix_step(H.build_list(atoms, box, 6.0))
    tag, sf, _, _ = o.slots()
    t = tag[10, 10, 10]
    assert list(t) == [7, 3], t
    assert sf[10, 10, 10, 0] > 0 and sf[10, 10, 10, 1] > 0
    # sphere 7 steps back (the node stays inside its zone, fraction 0): its
    # slot empties first in list order, and sphere 5 then takes it
    atoms.x[0] = node + np.array([r + 1.4, 0, 0])
    o.fix_step(H.build_list(atoms, box, 6.0))
    tag, _, _, _ = o.slots()
    assert list(tag[10, 10, 10]) == [5, 3]
    # a particle that jumps more than its zone margin leaves its slots behind:
    # only nodes inside the zone are visited (:59-63)
    atoms.x[2] = node + np.array([-8.0, 0, 0])
    o.fix_step(H.build_list(atoms, box, 6.0))
    tag, _, _, _ = o.slots()
    assert list(tag[10, 10, 10]) == [5, 3]


def test_held_slots_always_exceed_the_set_threshold():
    """The slot table of src/LBM_PSM_exchangeParticleData.cpp:81-106 only
    ever sets a slot with sf > 1e-5 and clears it to tag 0, sf 0.  After
    every step of random overlapping spheres that move, jump, vanish from the
    list and return, a held slot has sf > 1e-5 and an empty one sf = 0: the
    old fraction then only says whether the tag holds a slot, which is what
    the GPU map relies on to skip reading it."""
    n = (16, 16, 16)
    box = lattice_box(3, n, (1, 1, 1))
    o = O.Oracle(3, n, (1, 1, 1), 0.8)
    rng = np.random.default_rng(404)
    m = 7
    pos = rng.uniform(0.0, 16.0, (m, 3))
    radius = rng.uniform(1.2, 3.5, m)
    held_seen = 0
    for step in range(40):
        pos = (pos + rng.uniform(-1.5, 1.5, (m, 3))) % 16.0
        keep = rng.uniform(size=m) > 0.2
        atoms = H.Atoms(tag=np.arange(1, m + 1, dtype=np.int32)[keep], x=pos[keep].copy(),
                        v=np.zeros((keep.sum(), 3)), omega=np.zeros((keep.sum(), 3)),
                        radius=radius[keep].copy())
        o.fix_step(H.build_list(atoms, box, 6.0))
        tag, sf, _, _ = o.slots()
        held = tag != 0
        assert np.all(sf[held] > 1e-5), step
        assert np.all(sf[~held] == 0.0), step
        held_seen += int(held.sum())
    assert held_seen > 1000


def test_hydro_known_answer():
    """One node force F at lever r: T = r x F, S = -sym(r F), summed with
    the force and torque factors (src/LBM_PSM_exchangeParticleData.cpp:223-240)."""
    n = (12, 12, 12)
    u = O.lattice_units()
    u.forceFactor = 2.0
    u.torqueFactor = 3.0
    o = O.Oracle(3, n, (1, 1, 1), 0.8, unit=u)
    tagp = O.LIB.lbo_slot_tag(o._h)
    Fp = O.LIB.lbo_slot_F(o._h)
    nh = o.nh
    # node (ref indices) 7, 6, 6 holds tag 4 in slot 1 with force (0, 1, -2)
    idx = (7 * nh[1] + 6) * nh[2] + 6
    tagp[2 * idx + 1] = 4
    Fp[6 * idx + 3], Fp[6 * idx + 4], Fp[6 * idx + 5] = 0.0, 1.0, -2.0
    fh = np.zeros(3)
    th = np.zeros(3)
    st = np.zeros(6)
    xp = np.array([5.0, 5.0, 5.0])  # -> reference coords 6,6,6; lever r = (1, 0, 0)
    d = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    O.LIB.lbo_hydro(o._h, C.byref(u), d(o.origin), 4, d(xp), 1.5, d(fh), d(th), d(st))
    assert np.array_equal(fh, [0.0, 2.0, -4.0])
    # T = r x F = (1,0,0) x (0,1,-2) = (0, 2, 1) times 3
    assert np.array_equal(th, [0.0, 6.0, 3.0])
    # S_ab = -0.5 (r_a F_b + r_b F_a) * 3: xx 0, yy 0, zz 0, xy -1.5, xz 3, yz 0
    assert np.array_equal(st, [0.0, 0.0, 0.0, -1.5, 3.0, 0.0])


def test_units_stokes_deck():
    """UnitConversion for verification/3D_stokesDragForce-tau0p7/in.lbmpsm:22
    (Nlc 11, lc 0.0005, rho 1000, nu 1e-4, Re 0.1, tau 0.7)."""
    u = O.units(1000.0, 1e-4, 0.0005, 0.1, 11, 0.7, 3)
    assert u.dx == pytest.approx(5e-5, rel=1e-15)
    assert u.Uc == pytest.approx(0.02, rel=1e-15)
    assert u.u_lb == pytest.approx(0.1 * (0.2 / 3) / 10, rel=1e-14)
    # force factor: rho dx^4 / dt^2 with dt = dt_d * tc
    dt = u.dt_d * u.tc
    assert u.forceFactor == pytest.approx(1000.0 * 5e-5 ** 4 / dt ** 2, rel=1e-12)
    assert u.torqueFactor == pytest.approx(u.forceFactor * 5e-5, rel=1e-12)


def _fill_slots(o, slots):
    """Every node (halo included) holds the same slots: [(tag, B, u_s)]."""
    nodes = o.nodes_h
    tag = np.ctypeslib.as_array(O.LIB.lbo_slot_tag(o._h), shape=(nodes * 2,))
    sf = np.ctypeslib.as_array(O.LIB.lbo_slot_sf(o._h), shape=(nodes * 2,))
    us = np.ctypeslib.as_array(O.LIB.lbo_slot_u(o._h), shape=(nodes * 6,))
    tag[:] = 0
    sf[:] = 0.0
    us[:] = 0.0
    for s, (t, B, u) in enumerate(slots):
        tag[s::2] = t
        sf[s::2] = B
        for d in range(3):
            us[3 * s + d::6] = u[d]


@pytest.mark.parametrize("tau", [0.6, 0.8, 1.0, 1.7])
@pytest.mark.parametrize("slots", [
    [(3, 0.35, (0.004, -0.002, 0.001))],
    [(3, 0.3, (0.004, -0.002, 0.001)), (8, 0.45, (-0.003, 0.001, 0.002))],
    [(3, 0.7, (0.004, -0.002, 0.001)), (8, 0.6, (-0.003, 0.001, 0.002))],   # B1+B2 > 1
], ids=["one-slot", "two-slots", "two-slots-renormalised"])
def test_psm_uniform_medium_darcy_velocity(tau, slots):
    """A uniformly partially-saturated medium under a body force.  With the
    collision of src/LBM_PSM_BGK_dynamics.h:60-113 (solid term
    Omega_s = feq(rho, u_s) - f + (1 - omega)(f - feq(rho, u)), weights B_s
    renormalised when B_1 + B_2 > 1) and Guo forcing with the half-force
    velocity shift, the momentum of a uniform node obeys
        j' = (1 - B) j + rho sum_s B_s u_s + F (1 - B/2),   B = sum_s B_s,
    so the fluid relaxes geometrically (factor 1 - B per step) to the Darcy
    state u = sum_s B_s u_s / B + F / (rho B), independent of tau, and each
    slot's node force is F_s = B_s rho (u - u_s) - B_s omega F / 2.  A
    collision with omega and 1 - omega exchanged, without the (1 - B)
    factor on the fluid part, or with the fluid velocity in the solid
    equilibrium gives another steady state."""
    F = np.array([2e-5, -1e-5, 5e-6])
    o = O.Oracle(3, (4, 5, 3), (1, 1, 1), tau, F_lbm=F)
    _fill_slots(o, slots)
    Bs = np.array([b for _, b, _ in slots])
    Us = np.array([u for _, _, u in slots])
    Bt = Bs.sum()
    if Bt > 1.0:
        Bs = Bs / Bt
        Bt = 1.0
    for _ in range(400):
        O.LIB.lbo_halo_exchange(o._h)
        O.LIB.lbo_collide_stream(o._h)
    # one more step with the node forces cleared, to read one step's force
    Fp = np.ctypeslib.as_array(O.LIB.lbo_slot_F(o._h), shape=(o.nodes_h * 6,))
    Fp[:] = 0.0
    O.LIB.lbo_halo_exchange(o._h)
    O.LIB.lbo_collide_stream(o._h)
    rho, u = o.macros()
    rho0 = float(rho.mean())
    assert np.allclose(rho, 1.0, rtol=0, atol=1e-13)
    expect = (Bs @ Us) / Bt + F / (rho0 * Bt)
    assert np.max(np.abs(u.reshape(-1, 3) - expect)) < 1e-13
    omega = 1.0 / tau
    _, _, _, Fh = o.slots()
    for s in range(len(slots)):
        Fs = Bs[s] * rho0 * (expect - Us[s]) - Bs[s] * omega * F / 2.0
        assert np.max(np.abs(Fh.reshape(-1, 6)[:, 3 * s:3 * s + 3] - Fs)) < 1e-13


def test_periodic_array_stokes_drag_hasimoto():
    """Stokes drag of a simple cubic array of spheres (one sphere of radius
    5 in a periodic 32^3 box, volume fraction c = 0.016), fixed, in a flow
    driven by a body force: K = F_sphere / (6 pi mu a U) against Hasimoto's
    series as extended by Sangani & Acrivos (1982),
    K = 1 / (1 - 1.7601 c^1/3 + c - 1.5593 c^2 + 3.9799 c^8/3 - 3.0734 c^10/3),
    U the mean velocity over the cell.  The fluid starts at the series'
    steady mean velocity, so the drag ratio settles within 2000 steps.
    PSM with B = solid fraction under-resolves a radius-5 sphere by a few
    per cent at tau 0.8 (measured: K 3.5 % below the series); the bound is
    5 %.  The collision with omega and 1 - omega exchanged in the solid
    term, or without the (1 - B) factor on the fluid part, diverges here;
    with the fluid velocity in the solid equilibrium the drag vanishes."""
    L, a, tau, F = 32, 5.0, 0.8, 1e-6
    nu = (tau - 0.5) / 3.0
    c = 4.0 * math.pi * a ** 3 / 3.0 / L ** 3
    x = c ** (1.0 / 3.0)
    Kh = 1.0 / (1 - 1.7601 * x + x ** 3 - 1.5593 * x ** 6 + 3.9799 * x ** 8 - 3.0734 * x ** 10)
    o = O.Oracle(3, (L, L, L), (1, 1, 1), tau, F_lbm=(F, 0.0, 0.0))
    o.initialise(1.0, (F * L ** 3 / (6 * math.pi * nu * a * Kh), 0.0, 0.0))
    box = H.Box(lo=np.zeros(3), prd=np.array([L, L, L], float), periodic=(1, 1, 1))
    c0 = L / 2 + 0.3
    atoms = H.Atoms(tag=np.array([1], np.int32), x=np.array([[c0, c0 - 0.1, c0 + 0.2]]),
                    v=np.zeros((1, 3)), omega=np.zeros((1, 3)), radius=np.array([a]))
    lst = H.build_list(atoms, box, a + 4.0)
    for _ in range(2000):
        Fp, T, _ = o.fix_step(lst)
    Fo, _ = H.reduce_to_owners(lst, Fp, T)
    _, u = o.macros()
    U = float(u[..., 0].mean())
    K = Fo[0, 0] / (6.0 * math.pi * nu * a * U)
    assert abs(K / Kh - 1.0) < 0.05, (K, Kh)
    assert abs(Fo[0, 1]) < 1e-3 * Fo[0, 0] and abs(Fo[0, 2]) < 1e-3 * Fo[0, 0]


def _halo_slot_mismatch(n, per, pos, radii, steps):
    """Largest difference, over the steps, between the slots of a periodic
    axis's halo planes and those of their interior counterparts (tags
    compared as a count of differing slots), in the oracle's halo'd lattice."""
    box = lattice_box(3, n, per)
    atoms = random_atoms(box, radii, seed=29, positions=np.array(pos, float), vmax=1e-2,
                         wmax=1e-3)
    ora = O.Oracle(3, n, per, 0.7)
    worst = 0.0
    for step in range(steps):
        H.move_linear(atoms, box, step, 1.0)
        lst = H.build_list(atoms, box, float(np.max(atoms.radius)) + 4.0)
        ora.fix_step(lst)
        nh = ora.nh
        sf = ora._view(O.LIB.lbo_slot_sf, ora.nodes_h * 2).reshape(*nh, 2)
        tg = ora._view(O.LIB.lbo_slot_tag, ora.nodes_h * 2, np.int32).reshape(*nh, 2)
        for ax in range(3):
            if not per[ax]:
                continue
            m = nh[ax]
            inner = tuple(slice(1, -1) if d != ax else slice(None) for d in range(3))
            a_sf = np.take(sf, [0, m - 1], axis=ax)[inner]
            b_sf = np.take(sf, [m - 2, 1], axis=ax)[inner]
            a_tg = np.take(tg, [0, m - 1], axis=ax)[inner]
            b_tg = np.take(tg, [m - 2, 1], axis=ax)[inner]
            worst = max(worst, float(np.abs(a_sf - b_sf).max()), float((a_tg != b_tg).sum()))
    return worst


def test_periodic_image_zones_and_halo_slots():
    """The reference keeps a halo on periodic axes and maps every list entry,
    ghost images included, onto it, clearing a slot when an entry of the
    same tag finds no overlap there (src/LBM_PSM_exchangeParticleData.cpp:
    24-112).  When the map zones of one sphere's own images overlap (here
    r = 3.2 across a 9-node periodic y extent, below 2r + 5), a later image
    clears, on one side of the seam only, a slot an earlier image set: the
    halo plane then ends the map with slots that differ from its periodic
    counterpart, and the reference streams from it.  A lattice that keeps no
    periodic halo, such as the library, streams from the counterpart instead.
    With extents above 2r + 5 the halo planes equal their counterparts."""
    pos = [[5.0, 4.4, 31.7], [4.6, 4.8, 2.1]]
    assert _halo_slot_mismatch((10, 9, 64), (1, 1, 1), pos, [3.2, 2.8], 1) > 0
    pos = [[7.0, 6.4, 31.7], [6.6, 6.8, 2.1]]
    assert _halo_slot_mismatch((14, 13, 64), (1, 1, 1), pos, [3.2, 2.8], 4) == 0
