"""FP32 / FP16 population storage against the FP64 path on the same inputs.

FP32 stores f - w*rho0 and computes in FP32; FP16 stores the same deviations
scaled by 2^12 and computes in FP32; node forces stay per node, per-particle
force/torque/stresslet and the mass are summed in FP64.  The FP64 path is
the parity form (checked against the oracle in test_gpu_parity.py); the
error of each reduced-precision result is measured against it.

Bars: FP32 drag and torque within 1e-5 (relative to the largest particle
force / torque), velocity field within 1e-5 of the largest velocity, mass
drift below 1e-6 per node over the run.  FP16: drag, torque and velocity
within 2e-3 (measured values are printed; the storage rounds each stored
deviation to 2^-11 of itself), mass drift below 1e-4 per node.
"""
import numpy as np
import pytest

import lbmpsm
from lbmpsm import harness as H
import scenarios as S
from scenarios import lattice_box, random_atoms

pytestmark = pytest.mark.gpu


def run(storage, steps=300):
    n = (64, 64, 64)
    per = (1, 0, 1)
    box = lattice_box(3, n, per)
    atoms = random_atoms(box, [8.0] * 12, seed=21, vmax=1e-3, wmax=1e-4, gap=2.0)
    lat = lbmpsm.Lattice(lbmpsm.lattice_units_config(3, n, per, 0.7, storage=storage,
                                                     F_lbm=(2e-6, 0, 0)))
    lat.set_bc("shear")
    for s in range(steps):
        H.move_linear(atoms, box, s, 1.0)
        lst = H.build_list(atoms, box, 12.0)
        lat.set_particles(lst.tag, lst.x, lst.v, lst.omega, lst.radius)
        lat.step(store_macros=(s == steps - 1))
        F, T, S = lat.hydro()
    rho, u = lat.macros()
    m = lat.mass()
    lat.close()
    return F, T, S, u, m


@pytest.fixture(scope="module")
def ref64():
    print("\n[case] precision reference FP64 64^3 shear 12 spheres 300 steps", flush=True)
    return run("f64")


@pytest.mark.parametrize("storage,tol", [("f32", 1e-5), ("f16", 2e-3)])
def test_reduced_storage_error(ref64, storage, tol):
    print(f"\n[case] precision {storage}", flush=True)
    F, T, S, u, m = run(storage)
    F0, T0, S0, u0, m0 = ref64
    eF = np.max(np.linalg.norm(F - F0, axis=1)) / np.max(np.linalg.norm(F0, axis=1))
    eT = np.max(np.linalg.norm(T - T0, axis=1)) / np.max(np.linalg.norm(T0, axis=1))
    eS = np.max(np.abs(S - S0)) / np.max(np.abs(S0))
    eu = np.max(np.abs(u - u0)) / np.max(np.abs(u0))
    em = abs(m - m0) / (64 * 64 * 64)
    print(f"  {storage}: drag {eF:.2e} torque {eT:.2e} stresslet {eS:.2e} velocity {eu:.2e} "
          f"mass/node {em:.2e}", flush=True)
    assert eF <= tol and eT <= tol and eu <= tol
    assert em <= (1e-6 if storage == "f32" else 1e-4)


def run_contacts(storage, steps=120):
    """Three pairs of spheres of radius 6 with centres 12.3 apart (along x,
    along y, diagonally): their rims share nodes, so some waves hold a node
    with two particles and run the particle form with both slots."""
    n = (48, 48, 48)
    per = (1, 1, 1)
    box = lattice_box(3, n, per)
    d = 12.3
    pos = np.array([[12.0, 12.0, 12.0], [12.0 + d, 12.0, 12.0],
                    [30.0, 28.0, 20.0], [30.0, 28.0 + d, 20.0],
                    [14.0, 34.0, 36.0], [14.0 + d / np.sqrt(2), 34.0 + d / np.sqrt(2), 36.0]])
    atoms = random_atoms(box, [6.0] * 6, seed=29, positions=pos, vmax=1e-3, wmax=2e-4)
    lat = lbmpsm.Lattice(lbmpsm.lattice_units_config(3, n, per, 0.8, storage=storage,
                                                     F_lbm=(3e-6, 0, 0)))
    for s in range(steps):
        H.move_linear(atoms, box, s, 1.0)
        lst = H.build_list(atoms, box, 10.0)
        lat.set_particles(lst.tag, lst.x, lst.v, lst.omega, lst.radius)
        lat.step()
        F, T, S = lat.hydro()
    tag, _, _, _ = lat.slots()
    shared = int(np.count_nonzero((tag[..., 0] > 0) & (tag[..., 1] > 0)))
    lat.close()
    return F, T, shared


@pytest.mark.parametrize("storage,tol", [("f32", 1e-5), ("f16", 2e-3)])
def test_reduced_storage_with_shared_nodes(storage, tol):
    print(f"\n[case] precision {storage} with shared nodes, 48^3, 3 sphere pairs 0.3 apart",
          flush=True)
    F0, T0, shared0 = run_contacts("f64")
    F, T, shared = run_contacts(storage)
    assert shared0 > 0 and shared == shared0, (shared0, shared)
    eF = np.max(np.linalg.norm(F - F0, axis=1)) / np.max(np.linalg.norm(F0, axis=1))
    eT = np.max(np.linalg.norm(T - T0, axis=1)) / np.max(np.linalg.norm(T0, axis=1))
    print(f"  {storage}: {shared} shared nodes, drag {eF:.2e} torque {eT:.2e}", flush=True)
    assert eF <= tol and eT <= tol


def run_axis_force(storage, F, steps=200):
    n = (32, 32, 32)
    per = (1, 1, 1)
    box = lattice_box(3, n, per)
    atoms = random_atoms(box, [5.0, 4.5, 5.5], seed=23, vmax=1e-3, wmax=1e-4, gap=2.0)
    lat = lbmpsm.Lattice(lbmpsm.lattice_units_config(3, n, per, 0.7, storage=storage, F_lbm=F))
    for s in range(steps):
        H.move_linear(atoms, box, s, 1.0)
        lst = H.build_list(atoms, box, 9.0)
        lat.set_particles(lst.tag, lst.x, lst.v, lst.omega, lst.radius)
        lat.step()
    Fh, T, _ = lat.hydro()
    lat.close()
    return Fh, T


@pytest.mark.parametrize("F", [(0.0, 3e-6, 0.0), (0.0, 0.0, 3e-6)])
def test_reduced_storage_axis_force(F):
    """The FP32 and FP16 collides formed for a body force along y or z (the
    x form runs in the other cases) against FP64: drag and torque of 3 moving
    spheres in a periodic 32^3 box after 200 steps, bars as above."""
    print(f"\n[case] precision, body force {F}", flush=True)
    F0, T0 = run_axis_force("f64", F)
    for storage, tol in (("f32", 1e-5), ("f16", 2e-3)):
        Fh, T = run_axis_force(storage, F)
        eF = np.max(np.linalg.norm(Fh - F0, axis=1)) / np.max(np.linalg.norm(F0, axis=1))
        eT = np.max(np.linalg.norm(T - T0, axis=1)) / np.max(np.linalg.norm(T0, axis=1))
        print(f"  {storage}: drag {eF:.2e} torque {eT:.2e}", flush=True)
        assert eF <= tol and eT <= tol


@pytest.mark.parametrize("dim", [2, 3])
def test_fp32_mass_kept_at_raised_density(dim):
    """FP32 storage keeps the mass of a periodic box whose density sits 0.1
    above the stored shift rho0 = 1 (as the bottom of a packed column under
    its pressure drop does).  Equilibrium of rho = 1.1 plus long-wave density
    and velocity modes (amplitudes 1e-2 and 2e-3, which outlive the 500 steps
    run here, so neighbouring nodes never hold the same values), no force, no
    particles.  Collisions conserve mass, so the total may change only by the
    rounding of the stored values, which then has no preferred sign (~1e-10
    per node).  Two roundings would give it one: lattice weights whose floats
    do not sum to 1 (the nearest floats sum to 1 + 1.5e-8 in D3Q19,
    1 + 7.4e-9 in D2Q9) add omega * eps * drho per node and step, 1.1e-6
    (5.3e-7) per node over this run; and the product of the D3Q19 rest
    weight with drho, rounded on its own, falls on thirds of the result's
    unit and is biased by -0.07 of it at drho = 0.1, 1.7e-7 per node here.
    The remaining systematic term, 6e-8 * 1.5 u^2 * omega per node and step
    from the float axis weight, is below 1e-9 here."""
    n = (48, 48, 48) if dim == 3 else (256, 256, 1)
    print(f"\n[case] FP32 mass at rho 1.1, D{dim}Q{19 if dim == 3 else 9} {n}, 500 steps",
          flush=True)
    x, y, z = np.meshgrid(*[2 * np.pi * np.arange(m) / m for m in n], indexing="ij")
    rho = 1.1 + 1e-2 * np.sin(x + 2 * y + 0.3) * np.cos(z - 0.7)
    u = np.zeros(n + (3,))
    u[..., 0] = 2e-3 * np.sin(y + 0.1) * np.cos(2 * z + 0.2)
    u[..., 1] = 2e-3 * np.cos(x - 0.4) * np.sin(z + 0.5)
    if dim == 3:
        u[..., 2] = 2e-3 * np.sin(2 * x + 0.6) * np.cos(y - 0.3)
    lat = lbmpsm.Lattice(lbmpsm.lattice_units_config(dim, n, (1, 1, 1), 0.7, storage="f32"))
    lat.set_populations(S.equilibrium(dim, rho, u))
    m0 = lat.mass()
    for _ in range(500):
        lat.step()
    drift = abs(lat.mass() - m0) / (n[0] * n[1] * n[2])
    lat.close()
    print(f"  mass drift per node {drift:.3e}", flush=True)
    assert drift < 2e-8


def run_2d(storage, steps=1500):
    """D2Q9 channel 240 x 120, `boundary f p`, `lbm-psm-bc xFlow` (Zou-He
    inlet and density outlet in x, velocity faces on the y planes), tau 0.7,
    body force along x, 10 moving disks of radius 5 to 8."""
    n = (240, 120, 1)
    per = (0, 1, 1)
    box = lattice_box(2, n, per)
    atoms = random_atoms(box, [5.0, 6.0, 7.0, 8.0, 5.5, 6.5, 7.5, 5.0, 6.0, 7.0], seed=41,
                         vmax=2e-3, wmax=2e-4, gap=2.0)
    lat = lbmpsm.Lattice(lbmpsm.lattice_units_config(2, n, per, 0.7, storage=storage,
                                                     F_lbm=(2e-6, 0, 0), u_lb=0.01))
    lat.set_bc("xFlow")
    for s in range(steps):
        H.move_linear(atoms, box, s, 1.0)
        lst = H.build_list(atoms, box, 12.0)
        lat.set_particles(lst.tag, lst.x, lst.v, lst.omega, lst.radius)
        lat.step(store_macros=(s == steps - 1))
    F, T, S = lat.hydro()
    Fo, To = H.reduce_to_owners(lst, F, T)
    _, u = lat.macros()
    m = lat.mass()
    lat.close()
    return Fo, To, u, m


@pytest.fixture(scope="module")
def ref64_2d():
    print("\n[case] precision reference FP64 D2Q9 240x120 xFlow, 10 moving disks, 1500 steps",
          flush=True)
    return run_2d("f64")


@pytest.mark.parametrize("storage,tol", [("f32", 1e-5), ("f16", 2e-3)])
def test_reduced_storage_2d_channel(ref64_2d, storage, tol):
    """The D2Q9 forms of the reduced-storage collide (fluid and particle
    waves, the 2D Zou-He faces on shifted populations) against FP64: disk
    drag and torque and the velocity field within tol of their largest
    values (bars as in 3D: 1e-5 for FP32, 2e-3 for FP16)."""
    print(f"\n[case] precision D2Q9 {storage}", flush=True)
    F0, T0, u0, m0 = ref64_2d
    F, T, u, m = run_2d(storage)
    eF = np.max(np.linalg.norm(F - F0, axis=1)) / np.max(np.linalg.norm(F0, axis=1))
    eT = np.max(np.abs(T[:, 2] - T0[:, 2])) / np.max(np.abs(T0[:, 2]))
    eu = np.max(np.abs(u - u0)) / np.max(np.abs(u0))
    print(f"  {storage}: drag {eF:.2e} torque {eT:.2e} velocity {eu:.2e} "
          f"mass {abs(m - m0) / m0:.2e}", flush=True)
    assert eF <= tol and eT <= tol and eu <= tol


def run_large(storage, lists):
    n = (256, 256, 256)
    lat = lbmpsm.Lattice(lbmpsm.lattice_units_config(3, n, (1, 1, 1), 0.65, storage=storage,
                                                     F_lbm=(5e-6, 0.0, 0.0)))
    for lst in lists:
        lat.set_particles(lst.tag, lst.x, lst.v, lst.omega, lst.radius)
        lat.step()
        F, T, _ = lat.hydro()
    lat.close()
    return F, T


@pytest.fixture(scope="module")
def large_lists():
    from lbmpsm import workloads as W
    box, atoms = W.dense_packing()
    rng = np.random.default_rng(11)
    atoms.v = rng.uniform(-1e-3, 1e-3, (atoms.n, 3))
    atoms.omega = rng.uniform(-1e-4, 1e-4, (atoms.n, 3))
    return W.atom_lists(box, atoms, 100, cutoff=14.0)


@pytest.fixture(scope="module")
def ref64_large(large_lists):
    print(f"\n[case] precision reference FP64 256^3 periodic, {large_lists[-1].n} entries, "
          f"100 steps", flush=True)
    return run_large("f64", large_lists)


@pytest.mark.slow
@pytest.mark.parametrize("storage,tol", [("f32", 1e-5), ("f16", 2e-3)])
def test_reduced_storage_large_lattice(ref64_large, large_lists, storage, tol):
    """256^3 periodic (262 144 row chunks: the size from which the reduced
    collides run two-wave workgroups, dealt to the XCDs in contiguous
    ranges), phi = 0.30 of spheres d ~ U[10, 20] moving with random
    velocities, Fx = 5e-6, 100 steps: drag and torque against FP64."""
    print(f"\n[case] precision {storage} 256^3", flush=True)
    F, T = run_large(storage, large_lists)
    F0, T0 = ref64_large
    eF = np.max(np.linalg.norm(F - F0, axis=1)) / np.max(np.linalg.norm(F0, axis=1))
    eT = np.max(np.linalg.norm(T - T0, axis=1)) / np.max(np.linalg.norm(T0, axis=1))
    print(f"  {storage}: drag {eF:.2e} torque {eT:.2e}", flush=True)
    assert eF <= tol and eT <= tol
