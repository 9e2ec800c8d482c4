"""`fix lbm-psm-restart` and `fix lbm-psm-vti` (C++ host side over the
library) in the reference's file formats.

* restart: the file written at step 8 holds the populations after LBM step 7
  in the reference's halo'd layout (src/fix_LBM_PSM_writeReadRestart.cpp:
  80-90): its interior and non-periodic halo equal the oracle's lattice bit
  for bit; a second run that reads it continues exactly as the first one;
* vti: the step-8 VTI file holds the step-7 fields in the reference's
  layout (src/fix_LBM_PSM_writeVTI.h:82-253): SolidFraction = min(sf0+sf1, 1),
  Density = rho * rho_fluid, Velocity = u * Uc/u_lb, x fastest, equal to the
  oracle's fields; the init file has the equilibrium density and zero
  velocity.
"""
import os
import re
import struct

import numpy as np
import pytest

import lbmpsm
import lbmpsm_oracle as O
from lammps_mock import Lammps, LammpsError
from lbmpsm import harness as H

RHO, NU, LC, RE, NLC, TAU = 1000.0, 1e-4, 0.0005, 0.1, 11, 0.7
FIX = f"1 all lbm-psm every 1 Nlc {NLC} lc {LC} rho {RHO} nu {NU} Re {RE} tau {TAU}"
L = 0.0016


def _box():
    return H.Box(lo=np.zeros(3), prd=np.array([L, L, L]), periodic=(1, 0, 1))


def _atoms():
    return H.Atoms(tag=np.array([3, 7], np.int32),
                   x=np.array([[0.0004, 0.0007, 0.0008], [0.0011, 0.0009, 0.0007]]),
                   v=np.array([[0.001, 0.0, 0.0], [0.004, -0.001, 0.002]]),
                   omega=np.array([[0.0, 0.0, 3.0], [1.0, -2.0, 0.5]]),
                   radius=np.array([0.00025, 0.0002]))


def _oracle(box):
    cfg = lbmpsm.config_from_fix(3, NLC, LC, RHO, NU, RE, TAU, (0, 0, 0), box.lo, box.prd,
                                 box.periodic)
    ora = O.Oracle(3, (cfg.nx, cfg.ny, cfg.nz), box.periodic, TAU,
                   unit=O.units(RHO, NU, LC, RE, NLC, TAU, 3), origin=box.lo)
    return ora, cfg


def test_io_fix_arguments():
    lmp = Lammps(_box())
    with pytest.raises(LammpsError, match="requires a fix lbm-psm"):
        lmp.fix("r all lbm-psm-restart every 4 write 1 read 0")
    lmp.fix(FIX)
    with pytest.raises(LammpsError, match="Illegal fix lbm-psm-restart command"):
        lmp.fix("r all lbm-psm-restart every 4 write 1")
    with pytest.raises(LammpsError, match="Illegal fix lbm-psm-restart command"):
        lmp.fix("r all lbm-psm-restart every 4 write 2 read 0")
    with pytest.raises(LammpsError, match="Illegal fix lbm-psm-vti command"):
        lmp.fix("v all lbm-psm-vti every 0")
    with pytest.raises(LammpsError, match="Illegal fix lbm-psm-vti command"):
        lmp.fix("v all lbm-psm-vti every 4 binary 3")
    lmp.fix("v all lbm-psm-vti every 4 binary 0 useDouble 1")


@pytest.mark.gpu
def test_restart_file_and_continuation(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    box = _box()
    print("\n[case] restart 32x33x32 shear, write at step 8, continue 4 steps", flush=True)
    a = Lammps(box)
    a.fix(FIX)
    a.fix("2 all lbm-psm-bc shear")
    a.fix("r all lbm-psm-restart every 8 write 1 read 0")
    a.init()
    dt = a.dt
    ora, cfg = _oracle(box)
    atoms = _atoms()
    lists = []
    for step in range(1, 13):
        H.move_linear(atoms, box, step, dt)
        lists.append(H.build_list(atoms, box, cutoff=0.0004))
    fa = []
    for step in range(1, 13):
        fa.append(a.step(lists[step - 1]))
        if step <= 7:
            ora.fix_step(lists[step - 1], bc_keyword=1)
        if step == 8:
            raw = open("LBM-PSM-restartFile-processor-0.bin", "rb").read()
    count = struct.unpack("<I", raw[:4])[0]
    f = np.frombuffer(raw[4:], dtype=np.float64)
    nh = ora.nh
    assert count == f.size == nh[0] * nh[1] * nh[2] * 19
    f = f.reshape(nh[0], nh[1], nh[2], 19)
    ref = ora.f_halo()
    # interior and the non-periodic (y) halo: the lattice's own values
    assert np.array_equal(f[1:-1, :, 1:-1], ref[1:-1, :, 1:-1])
    # periodic halos: the periodic image of the interior
    assert np.array_equal(f[0, 1:-1, 1:-1], f[-2, 1:-1, 1:-1])
    assert np.array_equal(f[:, :, 0], f[:, :, -2])

    b = Lammps(box)
    b.fix(FIX)
    b.fix("2 all lbm-psm-bc shear")
    b.fix("r all lbm-psm-restart every 100 write 0 read 1")
    b.init()
    for s in range(8, 13):
        fb, tb, vb = b.step(lists[s - 1])
        assert np.array_equal(fb, fa[s - 1][0]) and np.array_equal(tb, fa[s - 1][1]), s
    a.close()
    b.close()
    ora.close()


def _read_vti(path):
    raw = open(path, "rb").read()
    head, _, tail = raw.partition(b'<AppendedData encoding="raw">_')
    text = head.decode()
    typ = np.float64 if 'type="Float64"' in text else np.float32
    arrays, off = [], 0
    for _ in range(3):
        nbytes = struct.unpack("<I", tail[off:off + 4])[0]
        arrays.append(np.frombuffer(tail[off + 4:off + 4 + nbytes], dtype=typ))
        off += 4 + nbytes
    return text, arrays


@pytest.mark.gpu
def test_vti_fields(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    box = _box()
    print("\n[case] vti 32x33x32 shear, output every 4 steps", flush=True)
    lmp = Lammps(box)
    lmp.fix(FIX)
    lmp.fix("2 all lbm-psm-bc shear")
    lmp.fix("v all lbm-psm-vti every 4 binary 1 useDouble 1")
    lmp.init()
    dt = lmp.dt
    ora, cfg = _oracle(box)
    atoms = _atoms()
    for step in range(1, 9):
        H.move_linear(atoms, box, step, dt)
        lst = H.build_list(atoms, box, cutoff=0.0004)
        lmp.step(lst)
        if step <= 7:
            ora.fix_step(lst, bc_keyword=1)
    folder = "flow_field_output_data"
    assert sorted(os.listdir(folder)) == sorted(
        [f"flow_field.{s:08d}{e}" for s in (0, 4, 8) for e in ("_rank0.vti", ".pvti")])
    nx, ny, nz = cfg.nx, cfg.ny, cfg.nz
    text, (B, rho, u) = _read_vti(os.path.join(folder, "flow_field.00000008_rank0.vti"))
    assert f'WholeExtent="0 {nx - 1} 0 {ny - 1} 0 {nz - 1}"' in text
    assert re.search(r'Spacing="5e-05 5e-05 5e-05"', text)
    r0, u0 = ora.macros()
    tag, sf, us, F = ora.slots()
    # x fastest in the file; oracle arrays are [x][y][z]
    to_file = lambda a: np.transpose(a, (2, 1, 0) + tuple(range(3, a.ndim))).reshape(-1, *a.shape[3:])
    Bref = np.minimum(sf.sum(axis=-1), 1.0)
    assert np.array_equal(B, to_file(Bref).ravel())
    assert np.array_equal(rho, (to_file(r0) * RHO).ravel())
    assert np.array_equal(u, (to_file(u0) * (cfg.Uc / cfg.u_lb)).ravel())
    _, (B0, rho0, v0) = _read_vti(os.path.join(folder, "flow_field.00000000_rank0.vti"))
    assert np.all(rho0 == RHO) and np.all(v0 == 0.0) and np.all(B0 == 0.0)
    lmp.close()
    ora.close()
