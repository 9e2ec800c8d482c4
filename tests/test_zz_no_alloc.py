"""lbmpsm_step_noalloc (the step that allocates nothing: static device
scratch for the reduction's group sums, results in caller device memory)
against lbmpsm_step + lbmpsm_get_hydro on the same inputs, through the C ABI.

Both run the same kernels, so populations, slots and per-entry force,
torque and stresslet must be bit-identical.  Two calls in a row, the second
with a shorter particle list, and the entry point's refusals (a macro flag
that would allocate, a null output)."""
import ctypes as C

import numpy as np
import pytest

import lbmpsm
from lbmpsm import _native as N
from lbmpsm import harness as H
from lbmpsm.devmem import DeviceArray
from scenarios import lattice_box, random_atoms

pytestmark = pytest.mark.gpu


def _atoms(box, k, seed):
    atoms = random_atoms(box, [4.0, 3.5, 4.5, 3.0, 3.8][:k], seed=seed, vmax=2e-2, wmax=3e-3,
                         gap=0.2)
    return atoms


def test_step_noalloc_matches_step():
    n = (40, 36, 32)
    per = (1, 0, 1)
    box = lattice_box(3, n, per)
    print("\n[case] step_noalloc vs step, 40x36x32 shear, 5 then 2 spheres", flush=True)
    cfg = lbmpsm.lattice_units_config(3, n, per, 0.8, F_lbm=(1e-5, 0.0, 0.0))
    ref = lbmpsm.Lattice(cfg)
    got = lbmpsm.Lattice(cfg)
    ref.set_bc("shear")
    got.set_bc("shear")
    for call, k in enumerate((5, 2)):
        atoms = _atoms(box, k, seed=7 + call)
        for step in range(3):
            H.move_linear(atoms, box, step, 1.0)
            lst = H.build_list(atoms, box, 8.5)
            for lat in (ref, got):
                lat.set_particles(lst.tag, lst.x, lst.v, lst.omega, lst.radius)
            ref.step()
            Fr, Tr, Sr = ref.hydro()
            out = DeviceArray(np.full((lst.n, 12), np.nan))
            rc = N.LIB.lbmpsm_step_noalloc(got._h, 0, C.c_void_p(out.address))
            assert rc == 0, N.LIB.lbmpsm_last_error(got._h)
            o = out.to_host()
            out.close()
            assert np.array_equal(o[:, 0:3], Fr), (call, step)
            assert np.array_equal(o[:, 3:6], Tr), (call, step)
            assert np.array_equal(o[:, 6:12], Sr), (call, step)
        assert np.abs(Fr).max() > 0.0
        assert np.array_equal(ref.populations(), got.populations())
        tr, sfr, usr, Fhr = ref.slots()
        tg, sfg, usg, Fhg = got.slots()
        assert np.array_equal(tr, tg) and np.array_equal(sfr, sfg)
        assert np.array_equal(usr, usg) and np.array_equal(Fhr, Fhg)
        print(f"  call {call}: {lst.n} entries, identical", flush=True)
    # refusals: the macro arrays would have to be allocated; null output
    out = DeviceArray(np.zeros((lst.n, 12)))
    assert N.LIB.lbmpsm_step_noalloc(got._h, N.STEP_STORE_MACROS,
                                     C.c_void_p(out.address)) == N.ERR_STATE
    assert N.LIB.lbmpsm_step_noalloc(got._h, 0, None) == N.ERR_ARG
    out.close()
    ref.close()
    got.close()
