"""The D2Q9 disk and the dense polydisperse D3Q19 workloads at full size.

* D2Q9 256^2, fully periodic, tau 0.65, one fixed disk d = 16 at the box
  centre, body force Fx = 1e-6, rest start, 5000 steps: the library against
  the CPU oracle, drag within 1e-12 of the sum of its terms' magnitudes and
  the velocity field within 1e-12 relative.
* D3Q19 fluidised-bed column 1024 x 256 x 256 (x vertical), xFlow: Zou-He
  velocity inlet U = 0.005 at x = 0 and density outlet, periodic y/z,
  gravity Fx = -2e-6, 20000 spheres d = 12 moving linearly, tau 0.7, FP64
  and FP32 storage for 10 000 steps: FP32 drag/torque within 1e-5 of FP64
  (measured 8.1e-7 / 9.3e-7), and the step rate of both forms.  The spheres are seeded over the whole
  column (phi = 0.27): 20000 of them in its lower third would need
  phi = 0.81, beyond random sequential addition.
* D3Q19 256^3 periodic, Fx = 5e-6, 30 % volume fraction of fixed spheres with
  d ~ U[10, 20] (seeded RSA, largest first), tau in {0.6, 0.65, 0.8}, 3000
  steps in FP64, FP32 and FP16 storage: per-sphere drag, torque and
  stresslet error of FP32 and FP16 against FP64, relative to the largest
  value.  FP32 within 1e-5 (measured 0.6e-6 to 3.6e-6: the flow here is
  driven by a body force of 5e-6 per node and step, so any rounding bias of
  the momentum a collision hands on shows in the drag; the FP32 equilibrium
  is formed to carry none), FP16 within 5e-3 (measured up to 2.0e-3).
"""
import numpy as np
import pytest

import lbmpsm
import lbmpsm_oracle as O
from lbmpsm import harness as H
from scenarios import lattice_box

pytestmark = pytest.mark.gpu


def test_disk_2d_256_5000_steps_against_oracle():
    print("\n[case] D2Q9 256^2 periodic disk d=16, Fx=1e-6, tau 0.65, 5000 steps", flush=True)
    n = (256, 256, 1)
    per = (1, 1, 1)
    box = lattice_box(2, n, per)
    atoms = H.Atoms(tag=np.array([1], np.int32), x=np.array([[128.0, 128.0, 0.0]]),
                    v=np.zeros((1, 3)), omega=np.zeros((1, 3)), radius=np.array([8.0]))
    lst = H.build_list(atoms, box, 12.0)
    F_lbm = (1e-6, 0.0, 0.0)
    lat = lbmpsm.Lattice(lbmpsm.lattice_units_config(2, n, per, 0.65, F_lbm=F_lbm))
    ora = O.Oracle(2, n, per, 0.65, F_lbm=F_lbm)
    for step in range(5000):
        last = step == 4999
        lat.set_particles(lst.tag, lst.x, lst.v, lst.omega, lst.radius)
        lat.step(store_macros=last)
        F, T, S = lat.hydro()
        Fo, To, So = ora.fix_step(lst)
    rho, u = lat.macros()
    rho_o, u_o = ora.macros()
    _, _, _, Fnode = ora.slots()
    scale = np.sum(np.abs(Fnode))  # every node force enters the disk's sum once
    print(f"  drag {F[0, 0]:.12e} (oracle {Fo[0, 0]:.12e})", flush=True)
    assert abs(F[0, 0] - Fo[0, 0]) <= 1e-12 * scale
    assert np.max(np.abs(F - Fo)) <= 1e-12 * scale
    assert np.max(np.abs(u - u_o)) <= 1e-12 * np.max(np.abs(u_o))
    assert np.max(np.abs(rho - rho_o)) <= 1e-14
    assert F[0, 0] > 0.0  # dragged along the body force
    lat.close()
    ora.close()


def _dense_packing():
    box = lattice_box(3, (256, 256, 256), (1, 1, 1))
    rng = np.random.default_rng(303)
    d = rng.uniform(10.0, 20.0, 4000)
    vol = np.pi / 6.0 * np.cumsum(d ** 3)
    n = int(np.searchsorted(vol, 0.30 * 256 ** 3)) + 1
    d = np.sort(d[:n])[::-1]
    pos = H.rsa_spheres(box, d / 2.0, 303, gap=0.0)
    atoms = H.Atoms(tag=np.arange(1, n + 1, dtype=np.int32), x=pos, v=np.zeros((n, 3)),
                    omega=np.zeros((n, 3)), radius=d / 2.0)
    return box, atoms


@pytest.mark.slow
@pytest.mark.parametrize("tau", [0.6, 0.65, 0.8])
def test_dense_polydisperse_reduced_storage(tau):
    box, atoms = _dense_packing()
    lst = H.build_list(atoms, box, 14.0)
    print(f"\n[case] D3Q19 256^3 phi=0.30, {atoms.n} spheres ({lst.n} entries), tau {tau}, "
          f"3000 steps x FP64/FP32/FP16", flush=True)
    res = {}
    for storage in ("f64", "f32", "f16"):
        lat = lbmpsm.Lattice(lbmpsm.lattice_units_config(3, (256, 256, 256), (1, 1, 1), tau,
                                                         F_lbm=(5e-6, 0.0, 0.0), storage=storage))
        for step in range(3000):
            lat.set_particles(lst.tag, lst.x, lst.v, lst.omega, lst.radius)
            lat.step()
        F, T, S = lat.hydro()
        Fo, To = H.reduce_to_owners(lst, F, T)
        res[storage] = (Fo, To, S)
        lat.close()
        print(f"  {storage}: mean drag {np.mean(Fo[:, 0]):.6e}", flush=True)
    F64, T64, S64 = res["f64"]
    for storage, bound in (("f32", 1e-5), ("f16", 5e-3)):
        F, T, S = res[storage]
        eF = np.max(np.linalg.norm(F - F64, axis=1)) / np.max(np.linalg.norm(F64, axis=1))
        eT = np.max(np.linalg.norm(T - T64, axis=1)) / np.max(np.linalg.norm(T64, axis=1))
        eS = np.max(np.abs(S - S64)) / np.max(np.abs(S64))
        print(f"  {storage} vs f64: drag {eF:.2e} torque {eT:.2e} stresslet {eS:.2e}", flush=True)
        assert eF <= bound and eT <= bound and eS <= bound


@pytest.mark.slow
def test_fluidised_bed_10000_steps_fp32_vs_fp64():
    """The fluidised-bed column as the configuration names it: 10 000 steps
    in FP64 and FP32 storage, 20 000 spheres d = 12 moving linearly (new
    positions mapped every step), xFlow inlet U = 0.005 / outlet rho = 1,
    gravity Fx = -2e-6, tau 0.7.  The FP32 drag and torque of every sphere
    within 1e-5 of FP64 (relative to the largest; measured 8.1e-7 and
    9.3e-7), the step rate of both.  The pressure drop over the bed holds
    the density 0.1 above the stored shift rho0 = 1 at the bottom, so the
    FP32 collision must keep mass exactly at drho != 0: with the nearest
    floats as lattice weights (summing to 1 + 1.5e-8) the column gains a
    spread mass source and every drag moves by 1.2e-5 (mean drag +5.6e-5)."""
    import time
    from lbmpsm import workloads as W
    c = W.BED
    box, atoms = W.fluidised_bed()
    steps = 10_000
    print(f"\n[case] D3Q19 1024x256x256 xFlow, {atoms.n} moving spheres d=12, tau 0.7, "
          f"{steps} steps x FP64/FP32", flush=True)
    res = {}
    for storage in ("f64", "f32"):
        lat = lbmpsm.Lattice(lbmpsm.lattice_units_config(3, c["n"], c["periodic"], c["tau"],
                                                         F_lbm=c["F"], storage=storage,
                                                         u_lb=c["u_inlet"]))
        lat.set_bc("xFlow")
        a = atoms.copy()
        t0 = time.perf_counter()
        for step in range(steps):
            H.move_linear(a, box, step, 1.0)
            lst = H.build_list(a, box, c["radius"] + 4.0)
            lat.set_particles(lst.tag, lst.x, lst.v, lst.omega, lst.radius)
            lat.step()
            if (step + 1) % 2500 == 0:
                print(f"  {storage} step {step + 1}", flush=True)
        F, T, S = lat.hydro()
        dt = (time.perf_counter() - t0) / steps
        Fo, To = H.reduce_to_owners(lst, F, T)
        res[storage] = (Fo, To)
        lat.close()
        print(f"  {storage}: {c['n'][0] * c['n'][1] * c['n'][2] / dt / 1e6:.0f} MLUPS "
              f"({dt * 1e3:.2f} ms/step, host list every step), mean drag "
              f"{np.mean(Fo[:, 0]):.6e}", flush=True)
        assert np.all(np.isfinite(Fo))
    F64, T64 = res["f64"]
    F32, T32 = res["f32"]
    eF = np.max(np.linalg.norm(F32 - F64, axis=1)) / np.max(np.linalg.norm(F64, axis=1))
    eT = np.max(np.linalg.norm(T32 - T64, axis=1)) / np.max(np.linalg.norm(T64, axis=1))
    print(f"  f32 vs f64: drag {eF:.2e} torque {eT:.2e}", flush=True)
    assert np.mean(F64[:, 0]) > 0.0   # the rising flow drags the bed upwards
    assert eF <= 1e-5 and eT <= 1e-5
