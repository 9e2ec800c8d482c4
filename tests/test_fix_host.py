"""The C++ `fix lbm-psm` / `fix lbm-psm-bc` host side (lammps-lbm-psm_amd/fix)
under a minimal LAMMPS time-step loop.

CPU: argument parsing and the errors of src/fix_LBM_PSM.cpp:15-83 and
src/fix_LBM_PSM_BC.cpp:14-127, the init-time checks and DEM time step of
src/fix_LBM_PSM.cpp:104-140, and the loud failure without a GPU.

GPU: a deck with `every 2`, periodic ghost images and the shear faces, run
through the fix and through the CPU oracle fed the same atom lists.  On LBM
steps each local atom's force and torque must be the sum of its own entry
and of its ghost images' (post_force + reverse_comm), the virial the sum of
every entry's stresslet; on the steps in between, the stored force and
torque of the atom's own entry (src/fix_LBM_PSM.cpp:183-264).
"""
import os

import numpy as np
import pytest

import lbmpsm
import lbmpsm_oracle as O
from lammps_mock import Lammps, LammpsError
from lbmpsm import harness as H
from scenarios import check_hydro, hydro_scales, owner_scales

pytestmark = pytest.mark.skipif(
    not os.path.exists(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                    "lammps-lbm-psm_amd", "fix", "libmock_lammps.so")),
    reason="lammps-lbm-psm_amd/fix not built")

# verification/3D_*-tau0p7/in.lbmpsm:8-22 parameters on a smaller box
RHO, NU, LC, RE, NLC, TAU = 1000.0, 1e-4, 0.0005, 0.1, 11, 0.7
FIX = f"1 all lbm-psm every 2 Nlc {NLC} lc {LC} rho {RHO} nu {NU} Re {RE} tau {TAU}"


def _box(L=0.0016, per=(1, 0, 1)):
    return H.Box(lo=np.zeros(3), prd=np.array([L, L, L]), periodic=per)


@pytest.mark.parametrize("args", [
    "1 all lbm-psm Nlc 11 lc 0.0005 rho 1000 nu 1e-4 Re 0.1",         # narg < 15
    FIX + " tau 0.5",                                                  # tau <= 0.5
    FIX + " bogus 1",                                                  # unknown keyword
    FIX.replace("every 2", "every 0"),
    FIX.replace("Nlc 11", "Nlc -3"),
    FIX.replace("rho 1000.0", "rho 0"),
    FIX + " Fext 1 2",                                                 # missing value
    FIX + " storage f8",
    FIX.replace("lc 0.0005", "lc 5e-4x"),                              # not a number
])
def test_fix_rejects_what_the_reference_rejects(args):
    lmp = Lammps(_box())
    with pytest.raises(LammpsError, match="Illegal fix lbm-psm command|Expected floating"):
        lmp.fix(args)


def test_bc_fix_keywords_and_order():
    lmp = Lammps(_box())
    with pytest.raises(LammpsError, match="requires a fix lbm-psm"):
        lmp.fix("2 all lbm-psm-bc shear")
    lmp.fix(FIX)
    for kw in ("shear", "xFlow", "channel", "channel-velIn-pressOut", "closedBox"):
        lmp.fix(f"b{kw} all lbm-psm-bc {kw}")
    with pytest.raises(LammpsError, match="Illegal fix lbm-psm-bc command"):
        lmp.fix("3 all lbm-psm-bc Shear")
    with pytest.raises(LammpsError, match="Illegal fix lbm-psm-bc command"):
        lmp.fix("3 all lbm-psm-bc")


def test_init_checks_mach_and_cell_width():
    lmp = Lammps(_box())
    lmp.fix(FIX.replace(f"Re {RE}", "Re 200"))  # u_lb = 200 nu_lb / 10 -> Ma > 0.3
    with pytest.raises(LammpsError, match="Mach"):
        lmp.init()
    lmp = Lammps(_box(L=0.00161))
    lmp.fix(FIX)
    with pytest.raises(LammpsError, match="cell width"):
        lmp.init()


def test_init_sets_dem_timestep_then_fails_loudly_without_gpu():
    from lbmpsm.devmem import device_count
    if device_count() > 0:
        pytest.skip("a GPU is present")
    lmp = Lammps(_box())
    lmp.fix(FIX)
    with pytest.raises(LammpsError, match="lattice creation failed"):
        lmp.init()
    u = O.units(RHO, NU, LC, RE, NLC, TAU, 3)
    assert lmp.dt == u.dt_d * u.tc / 2.0  # phys_time(1) / nevery, bit for bit


def _deck_atoms():
    # sphere 1 straddles the periodic x boundary (one ghost image), sphere 2
    # moves across the box
    return H.Atoms(tag=np.array([3, 7], np.int32),
                   x=np.array([[0.00008, 0.0007, 0.0008], [0.0009, 0.0009, 0.0007]]),
                   v=np.array([[0.0, 0.0, 0.0], [0.004, -0.001, 0.002]]),
                   omega=np.array([[0.0, 0.0, 3.0], [1.0, -2.0, 0.5]]),
                   radius=np.array([0.00025, 0.0002]))


@pytest.mark.gpu
def test_fix_deck_against_oracle():
    box = _box()
    lmp = Lammps(box)
    lmp.fix(FIX)
    lmp.fix("2 all lbm-psm-bc shear")
    lmp.init()
    dt = lmp.dt
    cfg = lbmpsm.config_from_fix(3, NLC, LC, RHO, NU, RE, TAU, (0, 0, 0), box.lo, box.prd,
                                 box.periodic)
    ora = O.Oracle(3, (cfg.nx, cfg.ny, cfg.nz), box.periodic, TAU,
                   unit=O.units(RHO, NU, LC, RE, NLC, TAU, 3), origin=box.lo)
    atoms = _deck_atoms()
    stored_F = np.zeros((atoms.n, 3))
    stored_T = np.zeros((atoms.n, 3))
    vir_ref = np.zeros(6)
    oF = oT = np.zeros(atoms.n)   # bounds of the sums of |terms| (none before the first LBM step)
    vs = np.zeros(1)
    print(f"\n[case] fix deck 32x33x32 every 2, 12 steps", flush=True)
    for step in range(1, 13):
        H.move_linear(atoms, box, step, dt)
        lst = H.build_list(atoms, box, cutoff=0.0004)
        assert lst.n > lst.nlocal  # the ghost image is in the list
        f, t, vir = lmp.step(lst)
        if step % 2 == 0:
            F, T, S = ora.fix_step(lst, bc_keyword=1)
            Fo, To = H.reduce_to_owners(lst, F, T)
            stored_F, stored_T = F[:lst.nlocal].copy(), T[:lst.nlocal].copy()
            vir_ref = S.sum(axis=0)
            u = ora.unit
            to, _, _, Fho = ora.slots()
            sF, sT = hydro_scales(lst, to, Fho, u.forceFactor, u.torqueFactor, u.dx)
            oF, oT = owner_scales(lst, sF, sT)
            vs = np.array([sT.sum()])
        else:
            Fo, To = stored_F, stored_T
        check_hydro(f, Fo, oF, 1e-12, f"step {step} f")
        check_hydro(t, To, oT, 1e-12, f"step {step} torque")
        check_hydro(vir[None, :], vir_ref[None, :], vs, 1e-12, f"step {step} virial")
    assert np.max(np.abs(f)) > 0.0
    ora.close()
    lmp.close()


@pytest.mark.gpu
def test_unfixed_bc_stops_applying_faces():
    """`unfix` of `fix lbm-psm-bc` between two runs: the second run (a new
    lattice, as every init builds one) has no Zou-He faces, as the reference
    stops applying them once its BC fix is gone."""
    box = _box()
    lmp = Lammps(box)
    lmp.fix(FIX)
    lmp.fix("2 all lbm-psm-bc shear")
    lmp.init()
    dt = lmp.dt
    atoms = _deck_atoms()
    print("\n[case] fix deck, shear run then unfix lbm-psm-bc and a second run", flush=True)
    for step in range(1, 5):
        H.move_linear(atoms, box, step, dt)
        lmp.step(H.build_list(atoms, box, cutoff=0.0004))
    lmp.unfix("2")
    lmp.init()
    cfg = lbmpsm.config_from_fix(3, NLC, LC, RHO, NU, RE, TAU, (0, 0, 0), box.lo, box.prd,
                                 box.periodic)
    ora = O.Oracle(3, (cfg.nx, cfg.ny, cfg.nz), box.periodic, TAU,
                   unit=O.units(RHO, NU, LC, RE, NLC, TAU, 3), origin=box.lo)
    for step in range(5, 11):
        H.move_linear(atoms, box, step, dt)
        lst = H.build_list(atoms, box, cutoff=0.0004)
        f, t, vir = lmp.step(lst)
        if step % 2 == 0:
            F, T, S = ora.fix_step(lst, bc_keyword=0)
            Fo, To = H.reduce_to_owners(lst, F, T)
            u = ora.unit
            to, _, _, Fho = ora.slots()
            oF, oT = owner_scales(lst, *hydro_scales(lst, to, Fho, u.forceFactor,
                                                     u.torqueFactor, u.dx))
            check_hydro(f, Fo, oF, 1e-12, f"step {step} f")
            check_hydro(t, To, oT, 1e-12, f"step {step} torque")
    ora.close()
    lmp.close()
