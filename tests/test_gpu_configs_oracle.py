"""The FP64 library against the CPU oracle on the full-size 3D workloads.

Same inputs on both sides (the mock-LAMMPS lists of every step, ghosts
included), a few steps each; the bar is the parity bar of the small cases:
slot tags and solid fractions bit-exact, populations, node velocities and
node forces within 1e-12 of their largest magnitude, and per-entry force,
torque and stresslet within 1e-12 relative per component (1e-12 of the sum
of the term magnitudes / 100 for cancellation-dominated components).

* D3Q19 192^3, `boundary p f p` with `lbm-psm-bc shear` (U = 0.01), tau 0.7,
  400 moving spheres d = 16 by seeded RSA (the throughput workload of the
  shear suspension), 12 steps.
* D3Q19 256^3 periodic, Fx = 5e-6, phi = 0.30 of fixed spheres with
  d ~ U[10, 20], tau 0.65, 4 steps.
* D3Q19 1024 x 256 x 256 fluidised-bed column (`xFlow`, inlet U = 0.005,
  outlet rho = 1, Fx = -2e-6), 20000 moving spheres d = 12, tau 0.7,
  3 steps.
* D3Q19 96^3 periodic, 27 spheres d = 12 on a jittered 3 x 3 x 3 grid (the
  adjoint workload), Fx = 5e-6, tau 0.7, 10 steps.
"""
import numpy as np
import pytest

import lbmpsm
import lbmpsm_oracle as O
from lbmpsm import harness as H
from lbmpsm import workloads as W
from scenarios import BC_IDS, check_hydro, hydro_scales

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

TOL = 1e-12


def _max_rel(a, b, chunk=1 << 24):
    """max |a - b| / max |b| without whole-array temporaries."""
    a = a.reshape(-1)
    b = b.reshape(-1)
    num = 0.0
    den = 0.0
    for s in range(0, a.size, chunk):
        x = a[s:s + chunk]
        y = b[s:s + chunk]
        num = max(num, float(np.max(np.abs(x - y))))
        den = max(den, float(np.max(np.abs(y))))
    return num / den if den > 0 else num


def _oracle_interior_f(ora):
    """Zero-copy view of the oracle's interior populations."""
    ptr = O.LIB.lbo_f(ora._h)
    f = np.ctypeslib.as_array(ptr, shape=(ora.nodes_h * ora.q,)).reshape(*ora.nh, ora.q)
    return f[1:-1, 1:-1, 1:-1]


def run_and_compare(n, per, tau, F, bc, u_lb, lists, label):
    print(f"\n[case] {label}: {n}, bc {bc}, tau {tau}, {len(lists)} steps, "
          f"{lists[-1].n} list entries", flush=True)
    lat = lbmpsm.Lattice(lbmpsm.lattice_units_config(3, n, per, tau, F_lbm=F, u_lb=u_lb))
    if bc != "none":
        lat.set_bc(bc)
    ora = O.Oracle(3, n, per, tau, F_lbm=F, unit=O.lattice_units(u_lb))
    for step, lst in enumerate(lists):
        Fo, To, So = ora.fix_step(lst, bc_keyword=BC_IDS[bc])
        lat.set_particles(lst.tag, lst.x, lst.v, lst.omega, lst.radius)
        lat.step()
        Fg, Tg, Sg = lat.hydro()
        print(f"  step {step} done", flush=True)
    tg, sfg, usg, Fhg = lat.slots()
    to, sfo, uso, Fho = ora.slots()
    assert np.array_equal(tg, to), "slot tags differ"
    assert np.array_equal(sfg, sfo), "solid fractions differ"
    e_us = _max_rel(usg, uso)
    e_Fh = _max_rel(Fhg, Fho)
    del usg, uso
    sF, sT = hydro_scales(lists[-1], to, Fho)
    del Fhg, Fho, sfg, sfo
    held = int(np.count_nonzero(to))
    e_pop = _max_rel(lat.populations(), _oracle_interior_f(ora))
    print(f"  {held} held slots; pop rel {e_pop:.2e}, slot-u rel {e_us:.2e}, "
          f"node-F rel {e_Fh:.2e}", flush=True)
    assert e_pop <= TOL and e_us <= TOL and e_Fh <= TOL
    check_hydro(Fg, Fo, sF, TOL, "F")
    check_hydro(Tg, To, sT, TOL, "T")
    check_hydro(Sg, So, sT, TOL, "S")
    assert np.max(np.abs(Fo)) > 0.0
    lat.close()
    ora.close()


def test_shear_suspension_192cube_against_oracle():
    box, atoms = W.shear_suspension()
    c = W.SHEAR
    lists = W.atom_lists(box, atoms, 12, cutoff=c["radius"] + 4.0)
    run_and_compare(c["n"], c["periodic"], c["tau"], (0.0, 0.0, 0.0), "shear", c["u_wall"], lists,
                    "192^3 shear, 400 moving spheres d=16")


def test_dense_polydisperse_256cube_against_oracle():
    box, atoms = W.dense_packing()
    lst = H.build_list(atoms, box, 14.0)
    run_and_compare((256, 256, 256), (1, 1, 1), 0.65, (5e-6, 0.0, 0.0), "none", 0.01, [lst] * 4,
                    f"256^3 periodic phi=0.30, {atoms.n} spheres d~U[10,20]")


def test_fluidised_bed_column_against_oracle():
    box, atoms = W.fluidised_bed()
    c = W.BED
    lists = W.atom_lists(box, atoms, 3, cutoff=c["radius"] + 4.0)
    run_and_compare(c["n"], c["periodic"], c["tau"], c["F"], "xFlow", c["u_inlet"], lists,
                    "1024x256x256 xFlow bed, 20000 moving spheres d=12")


def test_adjoint_grid_96cube_against_oracle():
    rng = np.random.default_rng(404)
    g = np.array([16.0, 48.0, 80.0])
    pos = np.array([[a, b, c] for a in g for b in g for c in g]) + rng.uniform(-3, 3, (27, 3))
    box = H.Box(lo=np.zeros(3), prd=np.array([96.0] * 3), periodic=(1, 1, 1), dim=3)
    atoms = H.Atoms(tag=np.arange(1, 28, dtype=np.int32), x=pos,
                    v=rng.uniform(-1e-3, 1e-3, (27, 3)), omega=rng.uniform(-1e-4, 1e-4, (27, 3)),
                    radius=np.full(27, 6.0))
    lst = H.build_list(atoms, box, 8.0)
    run_and_compare((96, 96, 96), (1, 1, 1), 0.7, (5e-6, 0.0, 0.0), "none", 0.01, [lst] * 10,
                    "96^3 periodic, 27 spheres d=12 on a jittered grid")
