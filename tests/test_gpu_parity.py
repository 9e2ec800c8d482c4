"""GPU library vs CPU oracle on identical mock-LAMMPS inputs (FP64 path).

Bar: particle-slot state (tags, solid fractions) bit-exact; populations,
rho/u and slot velocities/forces equal to the oracle within 1e-12 of the
largest magnitude (in practice bit-exact: the kernels keep the reference's
operand order); per-entry force/torque/stresslet within 1e-12 of the sum of
absolute terms (the reductions run in a different order).
"""
import numpy as np
import pytest

import lbmpsm
import lbmpsm_oracle as O
from lbmpsm import harness as H
from scenarios import BC_IDS, check_hydro, hydro_scales, lattice_box, random_atoms, rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-12


def run_pair(dim, n, periodic, tau, F, bc, atoms, steps, cutoff=None, move=True,
             store_macros=True, label=""):
    print(f"\n[case] {label} dim={dim} n={n} periodic={periodic} bc={bc} steps={steps}", flush=True)
    box = lattice_box(dim, n, periodic)
    lat = lbmpsm.Lattice(lbmpsm.lattice_units_config(dim, n, periodic, tau, F_lbm=F))
    ora = O.Oracle(dim, n, periodic, tau, F_lbm=F)
    if bc != "none":
        lat.set_bc(bc)
    out = []
    for step in range(steps):
        lst = None
        if atoms is not None:
            if move:
                H.move_linear(atoms, box, step, 1.0)
            cut = cutoff if cutoff is not None else float(np.max(atoms.radius)) + 4.0
            lst = H.build_list(atoms, box, cut)
        else:
            lst = H.AtomList(tag=np.zeros(0, np.int32), x=np.zeros((0, 3)), v=np.zeros((0, 3)),
                             omega=np.zeros((0, 3)), radius=np.zeros(0), owner=np.zeros(0, int),
                             nlocal=0)
        Fo, To, So = ora.fix_step(lst, bc_keyword=BC_IDS[bc])
        lat.set_particles(lst.tag, lst.x, lst.v, lst.omega, lst.radius)
        lat.step(store_macros=store_macros)
        Fg, Tg, Sg = lat.hydro()
        out.append((lst, (Fg, Tg, Sg), (Fo, To, So)))
    return lat, ora, out


def compare(lat, ora, out, check_macros=True):
    fg, fo = lat.populations(), ora.populations()
    assert fg.shape == fo.shape
    e_pop = rel_err(fg, fo)
    tg, sfg, usg, Fhg = lat.slots()
    to, sfo, uso, Fho = ora.slots()
    assert np.array_equal(tg, to), "slot tags differ"
    assert np.array_equal(sfg, sfo), "solid fractions differ"
    e_us = rel_err(usg, uso) if np.any(uso) else 0.0
    e_Fh = rel_err(Fhg, Fho) if np.any(Fho) else 0.0
    print(f"  pop rel {e_pop:.2e}  slot-u rel {e_us:.2e}  node-F rel {e_Fh:.2e}", flush=True)
    assert e_pop <= TOL
    assert e_us <= TOL
    assert e_Fh <= TOL
    if check_macros:
        rg, ug = lat.macros()
        ro, uo = ora.macros()
        assert rel_err(rg, ro) <= TOL
        assert rel_err(ug, uo) <= TOL or np.max(np.abs(ug - uo)) <= 1e-15
    lst, (Fg, Tg, Sg), (Fo, To, So) = out[-1]
    if lst.n:
        sF, sT = hydro_scales(lst, to, Fho)
        check_hydro(Fg, Fo, sF, TOL, "F")
        check_hydro(Tg, To, sT, TOL, "T")
        check_hydro(Sg, So, sT, TOL, "S")


def test_fluid_periodic_guo_3d():
    lat, ora, out = run_pair(3, (20, 18, 16), (1, 1, 1), 0.7, (1e-5, -2e-6, 3e-6), "none", None,
                             12, label="fluid")
    compare(lat, ora, out)
    fg, fo = lat.populations(), ora.populations()
    assert np.array_equal(fg, fo), "FP64 fluid populations are not bit-exact"


def test_fluid_periodic_guo_2d():
    lat, ora, out = run_pair(2, (32, 24, 1), (1, 1, 1), 0.65, (1e-5, 2e-6, 0.0), "none", None,
                             15, label="fluid2d")
    compare(lat, ora, out)
    assert np.array_equal(lat.populations(), ora.populations())


@pytest.mark.parametrize("dim,F", [(3, (0.0, 2e-6, 0.0)), (3, (0.0, 0.0, -3e-6)),
                                   (2, (0.0, 2e-6, 0.0))])
def test_body_force_along_one_axis(dim, F):
    """The collides formed for a body force along a single lattice axis (the
    y and z kernels; x runs in most other cases) with moving spheres, two of
    them sharing rim nodes: populations, slots and F/T/S against the oracle,
    and fluid-only steps bit for bit."""
    n = (24, 22, 20) if dim == 3 else (40, 36, 1)
    per = (1, 1, 1)
    box = lattice_box(dim, n, per)
    if dim == 3:
        pos = np.array([[8.0, 10.2, 9.9], [15.6, 10.4, 10.1], [18.5, 3.3, 15.2]])
        radii = [3.7, 3.6, 3.2]
    else:
        pos = np.array([[12.0, 14.2, 0.0], [21.6, 14.4, 0.0], [33.5, 5.3, 0.0]])
        radii = [4.7, 4.6, 4.2]
    atoms = random_atoms(box, radii, seed=17, positions=pos, vmax=1e-2, wmax=1e-3)
    lat, ora, out = run_pair(dim, n, per, 0.7, F, "none", atoms, 8, label=f"axis force {F}")
    compare(lat, ora, out)
    lat2, ora2, _ = run_pair(dim, n, per, 0.7, F, "none", None, 6, label=f"axis force fluid {F}")
    assert np.array_equal(lat2.populations(), ora2.populations())


@pytest.mark.parametrize("dim,n,per,F,bc,pos,radii", [
    (3, (14, 13, 128), (1, 1, 1), (1e-5, -2e-6, 3e-6), "none",
     [[7.0, 6.4, 63.7], [6.6, 6.8, 2.1], [7.2, 6.0, 100.3]], [3.2, 2.8, 3.0]),
    (3, (13, 13, 62), (1, 1, 0), (0.0, 2e-6, 0.0), "none",
     [[6.5, 6.2, 31.5], [6.0, 5.9, 8.0], [6.8, 6.4, 52.2]], [3.3, 3.0, 2.9]),
    (3, (14, 14, 64), (1, 0, 1), (0.0, 0.0, 0.0), "shear",
     [[7.0, 7.1, 63.2], [6.8, 6.9, 30.0]], [3.5, 3.2]),
    (2, (20, 128, 1), (1, 1, 1), (1e-5, 2e-6, 0.0), "none",
     [[10.0, 63.6, 0.0], [9.4, 12.0, 0.0], [10.5, 120.4, 0.0]], [4.5, 4.0, 4.2]),
    (2, (24, 62, 1), (1, 0, 1), (0.0, 0.0, 0.0), "shear",
     [[12.0, 31.0, 0.0], [8.0, 10.3, 0.0]], [5.0, 4.5]),
])
def test_rows_of_whole_blocks(dim, n, per, F, bc, pos, radii):
    """Stored rows (z in 3D, y in 2D) a whole number of 64-node blocks long
    (periodic 64 or 128 nodes, or 62 interior nodes plus the two wall
    layers) take the collide's split addressing, a per-row scalar base plus
    a per-column lane offset.  Periodic and walled row ends, rows of two
    blocks whose pulls cross the block boundary, spheres on that boundary
    and on the periodic seam, general and single-axis body forces: against
    the oracle.  Each periodic extent exceeds 2r + 5 nodes, so the map
    zones of a sphere's own images stay apart (with them overlapping, the
    reference's halo nodes can end a map with slots that differ from their
    periodic counterparts, which this library does not store)."""
    box = lattice_box(dim, n, per)
    atoms = random_atoms(box, radii, seed=29, positions=np.array(pos, float), vmax=1e-2,
                         wmax=1e-3)
    lat, ora, out = run_pair(dim, n, per, 0.7, F, bc, atoms, 8, label=f"whole blocks {n}")
    compare(lat, ora, out)


def test_spheres_shear_3d_moving():
    n = (28, 26, 24)
    per = (1, 0, 1)
    box = lattice_box(3, n, per)
    atoms = random_atoms(box, [4.0, 3.5, 4.5, 3.0], seed=3, vmax=2e-2, wmax=3e-3, gap=0.2)
    lat, ora, out = run_pair(3, n, per, 0.8, (0.0, 0.0, 0.0), "shear", atoms, 10, label="shear")
    compare(lat, ora, out)


def test_spheres_periodic_guo_shared_nodes():
    # two spheres nearly touching (shared rim nodes) and one crossing x = 0
    n = (24, 24, 24)
    per = (1, 1, 1)
    pos = np.array([[10.0, 12.2, 11.9], [17.6, 12.4, 12.1], [0.7, 5.3, 18.2]])
    box = lattice_box(3, n, per)
    atoms = random_atoms(box, [3.7, 3.6, 3.2], seed=5, positions=pos, vmax=1e-2, wmax=1e-3)
    lat, ora, out = run_pair(3, n, per, 0.65, (2e-6, 0.0, 0.0), "none", atoms, 8, label="shared")
    compare(lat, ora, out)
    tag, sf, _, _ = lat.slots()
    assert np.any((tag[..., 0] > 0) & (tag[..., 1] > 0)), "scenario has no shared node"


def test_three_particles_one_node_drop():
    # three spheres whose rims meet near one node: the third is dropped there
    n = (20, 20, 20)
    per = (1, 1, 1)
    c = np.array([10.0, 10.0, 10.0])
    r = 3.0
    d = r + 0.55
    pos = np.array([c + [d, 0, 0], c + [-d / 2, d * 0.866, 0], c + [-d / 2, -d * 0.866, 0]])
    box = lattice_box(3, n, per)
    atoms = random_atoms(box, [r, r, r], seed=1, positions=pos, vmax=0.0, wmax=0.0)
    lat, ora, out = run_pair(3, n, per, 0.9, (0, 0, 0), "none", atoms, 3, move=False,
                             label="drop")
    compare(lat, ora, out)


def test_particle_leaves_list_and_returns():
    # a sphere missing from the list for some steps leaves stale slots: they
    # keep their state and their node forces keep accumulating; when it comes
    # back its slots are set again (force restarted from zero)
    n = (24, 22, 20)
    per = (1, 1, 1)
    box = lattice_box(3, n, per)
    atoms = random_atoms(box, [3.2, 2.8, 3.0], seed=17, vmax=8e-3, wmax=1e-3, gap=1.5)
    lat = lbmpsm.Lattice(lbmpsm.lattice_units_config(3, n, per, 0.7, F_lbm=(1e-6, 0, 0)))
    ora = O.Oracle(3, n, per, 0.7, F_lbm=(1e-6, 0, 0))
    print("\n[case] particle leaves the list for steps 4-7 and returns", flush=True)
    out = []
    for step in range(12):
        H.move_linear(atoms, box, step, 1.0)
        if 4 <= step < 8:  # tag 2 (and its images) is missing from the list
            k = atoms.tag != 2
            sub = H.Atoms(tag=atoms.tag[k], x=atoms.x[k], v=atoms.v[k], omega=atoms.omega[k],
                          radius=atoms.radius[k])
            lst = H.build_list(sub, box, 7.5)
        else:
            lst = H.build_list(atoms, box, 7.5)
        Fo, To, So = ora.fix_step(lst, bc_keyword=0)
        lat.set_particles(lst.tag, lst.x, lst.v, lst.omega, lst.radius)
        lat.step(store_macros=True)
        Fg, Tg, Sg = lat.hydro()
        out.append((lst, (Fg, Tg, Sg), (Fo, To, So)))
        if step == 7:
            tag, _, _, Fh = lat.slots()
            s0, s1 = tag[..., 0] == 2, tag[..., 1] == 2
            assert np.any(s0 | s1), "no stale slot of the missing sphere"
            assert np.any(Fh[..., 0:3][s0] != 0.0) or np.any(Fh[..., 3:6][s1] != 0.0), \
                "stale slots carry no force"
            compare(lat, ora, out)
    compare(lat, ora, out)


def test_slots_read_between_map_and_collide():
    # the map's slot state is visible before the step: tags, solid fractions
    # and node velocities as the oracle leaves them (its collide does not
    # change them), and zero force on every slot the map set, as the
    # reference's set resets it (src/LBM_PSM_lattice.cpp:257-299)
    n = (20, 18, 22)
    per = (1, 1, 1)
    box = lattice_box(3, n, per)
    atoms = random_atoms(box, [3.1, 2.7, 3.4, 2.9], seed=23, vmax=2e-2, wmax=2e-3, gap=1.0)
    lat = lbmpsm.Lattice(lbmpsm.lattice_units_config(3, n, per, 0.8, F_lbm=(2e-6, 0, 0)))
    ora = O.Oracle(3, n, per, 0.8, F_lbm=(2e-6, 0, 0))
    print("\n[case] slots after set_particles, before step, 6 steps", flush=True)
    out = []
    for step in range(6):
        H.move_linear(atoms, box, step, 1.0)
        lst = H.build_list(atoms, box, 7.5)
        Fo, To, So = ora.fix_step(lst, bc_keyword=0)
        lat.set_particles(lst.tag, lst.x, lst.v, lst.omega, lst.radius)
        tg, sfg, usg, Fhg = lat.slots()
        to, sfo, uso, _ = ora.slots()
        assert np.array_equal(tg, to), f"step {step}: slot tags differ"
        assert np.array_equal(sfg, sfo), f"step {step}: solid fractions differ"
        assert rel_err(usg, uso) <= TOL, f"step {step}: node velocities differ"
        assert not np.any(Fhg), f"step {step}: a slot set by the map shows a force"
        lat.step(store_macros=True)
        Fg, Tg, Sg = lat.hydro()
        out.append((lst, (Fg, Tg, Sg), (Fo, To, So)))
    compare(lat, ora, out)


@pytest.mark.parametrize("L", [12, 16, 17])
def test_images_near_their_own_zones(L):
    """A sphere of radius 5 straddling a periodic edge: its images lie a box
    length away, inside each other's map and reduction zones for L = 12 and
    16 and just clear of them for L = 17.  Every step's force, torque and
    stresslet match the oracle, as do the slots at the end."""
    n = (L, L, L)
    print(f"\n[case] images L={L}", flush=True)
    box = lattice_box(3, n, (1, 1, 1))
    atoms = H.Atoms(tag=np.array([6, 2], np.int32),
                    x=np.array([[0.7, L - 1.2, 0.4], [L / 2.0 + 0.3, L / 2.0, L / 2.0 - 0.2]]),
                    v=np.array([[0.03, -0.02, 0.025], [-0.01, 0.0, 0.005]]),
                    omega=np.array([[1e-3, -2e-3, 5e-4], [0.0, 1e-3, 0.0]]),
                    radius=np.array([5.0, 2.5]))
    F = (2e-6, 0.0, 0.0)
    lat = lbmpsm.Lattice(lbmpsm.lattice_units_config(3, n, (1, 1, 1), 0.7, F_lbm=F))
    ora = O.Oracle(3, n, (1, 1, 1), 0.7, F_lbm=F)
    out = []
    for step in range(8):
        H.move_linear(atoms, box, step, 10.0)  # 0.2-0.4 nodes per step
        lst = H.build_list(atoms, box, 9.0)
        Fo, To, So = ora.fix_step(lst)
        lat.set_particles(lst.tag, lst.x, lst.v, lst.omega, lst.radius)
        lat.step()
        Fg, Tg, Sg = lat.hydro()
        to, _, _, Fho = ora.slots()
        sF, sT = hydro_scales(lst, to, Fho)
        check_hydro(Fg, Fo, sF, TOL, f"step {step} F")
        check_hydro(Tg, To, sT, TOL, f"step {step} T")
        check_hydro(Sg, So, sT, TOL, f"step {step} S")
        out.append((lst, (Fg, Tg, Sg), (Fo, To, So)))
    assert out[-1][0].n > 2  # the edge sphere has images
    compare(lat, ora, out, check_macros=False)
    lat.close()


def test_crowded_bricks_overflow_their_lists():
    # 7x7x9 tiny spheres 1.2 apart: more entries touch one 4x4x16-node brick
    # than its list holds, so the map replays the whole particle list there
    n = (24, 24, 32)
    per = (1, 1, 1)
    g = np.arange(7) * 1.2 + 6.3
    gz = np.arange(9) * 1.2 + 9.1
    pos = np.array([[x, y, z] for x in g for y in g for z in gz])
    box = lattice_box(3, n, per)
    atoms = random_atoms(box, [0.45] * len(pos), seed=9, positions=pos, vmax=2e-2, wmax=1e-2)
    # entries whose (narrowed) zone reaches the brick around (8, 8, 16)
    lo = np.array([8, 8, 0]) // [4, 4, 16] * [4, 4, 16]
    near = np.all((pos + 1.45 >= lo) & (pos - 1.45 < lo + [4, 4, 16]), axis=1)
    assert near.sum() > 64
    lat, ora, out = run_pair(3, n, per, 0.8, (1e-6, 0.0, 0.0), "none", atoms, 3,
                             cutoff=2.0, label="crowded")
    compare(lat, ora, out)


@pytest.mark.parametrize("bc", ["shear", "xFlow", "channel", "channel-velIn-pressOut",
                                "closedBox"])
@pytest.mark.parametrize("per", [(0, 0, 0), (1, 0, 1), (1, 1, 1), (0, 1, 1)])
def test_bc_keywords_3d(bc, per):
    n = (18, 16, 14)
    box = lattice_box(3, n, per)
    atoms = random_atoms(box, [3.0, 2.5], seed=11, vmax=5e-3, wmax=5e-4, gap=1.5)
    lat, ora, out = run_pair(3, n, per, 0.75, (1e-6, 0, 0), bc, atoms, 6,
                             label=f"bc3d-{bc}-{per}")
    compare(lat, ora, out)


@pytest.mark.parametrize("bc", ["shear", "xFlow", "channel", "channel-velIn-pressOut",
                                "closedBox"])
@pytest.mark.parametrize("per", [(0, 0, 1), (1, 0, 1), (1, 1, 1)])
def test_bc_keywords_2d(bc, per):
    n = (30, 22, 1)
    box = lattice_box(2, n, per)
    atoms = random_atoms(box, [4.0, 3.0], seed=7, vmax=5e-3, wmax=5e-4, gap=1.5)
    lat, ora, out = run_pair(2, n, per, 0.7, (1e-6, 0, 0), bc, atoms, 8, label=f"bc2d-{bc}-{per}")
    compare(lat, ora, out)


def test_disk_2d_fixed_guo():
    # config-1 geometry at reduced size: disk d=16 at the centre, Fx = 1e-6
    n = (64, 64, 1)
    per = (1, 1, 1)
    box = lattice_box(2, n, per)
    atoms = random_atoms(box, [8.0], seed=0, positions=np.array([[32.0, 32.0, 0.0]]), vmax=0.0,
                         wmax=0.0)
    lat, ora, out = run_pair(2, n, per, 0.65, (1e-6, 0, 0), "none", atoms, 30, move=False,
                             label="disk2d")
    compare(lat, ora, out)


def test_populations_roundtrip_and_mass():
    n = (12, 10, 8)
    per = (1, 0, 1)
    print("\n[case] set/get populations, mass, continue 4 shear steps", flush=True)
    lat = lbmpsm.Lattice(lbmpsm.lattice_units_config(3, n, per, 0.7))
    ora = O.Oracle(3, n, per, 0.7)
    rng = np.random.default_rng(2)
    f = ora.populations() * (1.0 + 0.01 * rng.standard_normal((12, 10, 8, 19)))
    lat.set_populations(f)
    assert np.array_equal(lat.populations(), f)
    assert abs(lat.mass() - f.sum()) <= 1e-12 * f.sum()
    fh = ora.f_halo()
    fh[1:-1, 1:-1, 1:-1] = f
    ora.set_f_halo(fh)
    lat.set_bc("shear")
    empty = H.AtomList(tag=np.zeros(0, np.int32), x=np.zeros((0, 3)), v=np.zeros((0, 3)),
                       omega=np.zeros((0, 3)), radius=np.zeros(0), owner=np.zeros(0, int),
                       nlocal=0)
    for _ in range(4):
        ora.fix_step(empty, bc_keyword=1)
        lat.set_particles(empty.tag, empty.x, empty.v, empty.omega, empty.radius)
        lat.step()
    assert rel_err(lat.populations(), ora.populations()) <= TOL


def test_zero_particles_and_errors():
    lat = lbmpsm.Lattice(lbmpsm.lattice_units_config(3, (8, 8, 8), (1, 1, 1), 0.7))
    lat.set_particles(np.zeros(0, np.int32), np.zeros((0, 3)), np.zeros((0, 3)), np.zeros((0, 3)),
                      np.zeros(0))
    lat.step()
    F, T, S = lat.hydro()
    assert F.shape == (0, 3)
    with pytest.raises(lbmpsm.LbmPsmError):
        lat.set_bc(99, 0.0)
    with pytest.raises(lbmpsm.LbmPsmError):
        lbmpsm.Lattice(lbmpsm.lattice_units_config(3, (8, 8, 8), (1, 1, 1), 0.5))


@pytest.mark.parametrize("per", [(1, 0, 1), (0, 0, 1)])
def test_moving_sphere_wall_axis_without_bc_long(per):
    """The Stokes-drag deck's setting at reduced size: `boundary p f p` and no
    lbm-psm-bc, so the y halo layers evolve on their own (stale values,
    src/LBM_PSM_MPICOMM.h:129,140), with a sphere crossing the periodic x
    boundary; 600 steps."""
    n = (30, 31, 30)
    box = lattice_box(3, n, per)
    pos = np.array([[27.5, 15.0, 15.0]])
    atoms = random_atoms(box, [3.0], seed=9, positions=pos, vmax=0.0, wmax=0.0)
    atoms.v[:] = [[0.01, 0.0, 0.0]]
    lat, ora, out = run_pair(3, n, per, 0.7, (0, 0, 0), "none", atoms, 600, cutoff=8.0,
                             label="stokes-like")
    compare(lat, ora, out)


def test_two_lists_set_before_one_step():
    """A list set twice before a step: the second list drops the sphere
    that shares nodes with another.  Slots the first map set keep a zero
    force (src/LBM_PSM_lattice.cpp:257-299 resets it when setting), also
    where the second map does not touch them."""
    n = (24, 24, 24)
    per = (1, 1, 1)
    box = lattice_box(3, n, per)
    pos = np.array([[9.0, 12.2, 11.9], [15.4, 12.4, 12.1], [3.1, 4.3, 18.2]])
    atoms = random_atoms(box, [3.6, 3.3, 2.8], seed=11, positions=pos, vmax=5e-3, wmax=1e-3)
    print("\n[case] two lists before one step, 24^3, spheres 1-2 share nodes", flush=True)
    lat = lbmpsm.Lattice(lbmpsm.lattice_units_config(3, n, per, 0.8, F_lbm=(1e-5, 0, 0)))
    ora = O.Oracle(3, n, per, 0.8, F_lbm=(1e-5, 0, 0))
    lst = H.build_list(atoms, box, 8.0)
    for step in range(4):  # slot forces build up
        ora.fix_step(lst)
        lat.set_particles(lst.tag, lst.x, lst.v, lst.omega, lst.radius)
        lat.step()
    tg, _, _, _ = lat.slots()
    shared = (tg[..., 0] == 1) & (tg[..., 1] == 2)
    assert shared.any(), "spheres 1 and 2 share no node"
    keep = atoms.tag != 1
    atoms_b = H.Atoms(tag=atoms.tag[keep], x=atoms.x[keep], v=atoms.v[keep],
                      omega=atoms.omega[keep], radius=atoms.radius[keep])
    lst_b = H.build_list(atoms_b, box, 8.0)
    ora.map(lst)
    Fo, To, So = ora.fix_step(lst_b)
    lat.set_particles(lst.tag, lst.x, lst.v, lst.omega, lst.radius)
    lat.set_particles(lst_b.tag, lst_b.x, lst_b.v, lst_b.omega, lst_b.radius)
    lat.step(store_macros=True)
    Fg, Tg, Sg = lat.hydro()
    compare(lat, ora, [(lst_b, (Fg, Tg, Sg), (Fo, To, So))])


def _rank_segments(n, p, s_lo, s_hi):
    """Ranges of one face on a rank grid: ranks split the n interior nodes
    of an axis n // p each, the last taking the remainder
    (src/LBM_PSM_lattice.cpp:34-63), and each applies the face over its own
    local range [1 + s_lo, n_local - s_hi] (src/fix_LBM_PSM_BC.cpp:223-271)."""
    out = []
    for r in range(p):
        ln = n // p + (n % p if r == p - 1 else 0)
        st = 1 + r * (n // p)
        out.append((st + s_lo, st + ln - 1 - s_hi))
    return out


@pytest.mark.parametrize("grid", [(3, 2, 1), (4, 1, 1)])
def test_xflow_2d_on_a_reference_rank_grid(grid):
    """lbmpsm_set_bc_ranks against the oracle with the same faces applied
    per rank through its face functions: 2D xFlow, whose y faces leave out
    the first and last interior node of every x-rank."""
    n = (38, 23, 1)
    per = (1, 1, 1)
    U = 0.01
    box = lattice_box(2, n, per)
    atoms = random_atoms(box, [4.0, 3.0], seed=21, vmax=5e-3, wmax=1e-3, gap=0.5)
    print(f"\n[case] 2D xFlow on rank grid {grid}, {n}", flush=True)
    lat = lbmpsm.Lattice(lbmpsm.lattice_units_config(2, n, per, 0.7, u_lb=U))
    lat.set_bc("xFlow", procgrid=grid)
    ora = O.Oracle(2, n, per, 0.7, unit=O.lattice_units(U))
    for step in range(12):
        H.move_linear(atoms, box, step, 1.0)
        lst = H.build_list(atoms, box, 8.0)
        ora.fix_step(lst, bc_keyword=0)
        for (a0, a1) in _rank_segments(n[1], grid[1], 0, 0):
            ora.zou_he_face(0, 1, (a0, a1), u=(U, 0, 0))             # xn
        for (a0, a1) in _rank_segments(n[1], grid[1], 0, 0):
            ora.zou_he_face(2, n[0], (a0, a1), rho=1.0)               # density xp
        for (a0, a1) in _rank_segments(n[0], grid[0], 1, 1):
            ora.zou_he_face(3, 1, (a0, a1), u=(U, 0, 0))             # yn
        for (a0, a1) in _rank_segments(n[0], grid[0], 1, 1):
            ora.zou_he_face(4, n[1], (a0, a1), u=(U, 0, 0))          # yp
        lat.set_particles(lst.tag, lst.x, lst.v, lst.omega, lst.radius)
        lat.step()
    assert np.array_equal(lat.populations(), ora.populations())
    lat.close()
