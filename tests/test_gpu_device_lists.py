"""lbmpsm_set_particles_device (list already in device memory) against
lbmpsm_set_particles (host list) on the same steps: slots, populations and
per-entry force/torque/stresslet must be identical, including when
max_radius understates a radius (those entries take the in-kernel path)."""
import numpy as np
import pytest

import lbmpsm
from lbmpsm import harness as H
from lbmpsm.devmem import DeviceArray
from scenarios import lattice_box, random_atoms

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("rmax_scale", [1.0, 0.5])
def test_device_list_equals_host_list(rmax_scale):
    print(f"\n[case] device-resident lists, 40^3 shear, 6 spheres, rmax x{rmax_scale}", flush=True)
    n = (40, 40, 40)
    per = (1, 0, 1)
    box = lattice_box(3, n, per)
    atoms = random_atoms(box, [5.0, 6.5, 4.0, 5.5, 3.0, 6.0], seed=5, gap=0.5)
    cfg = lbmpsm.lattice_units_config(3, n, per, 0.7)
    a, b = lbmpsm.Lattice(cfg), lbmpsm.Lattice(cfg)
    a.set_bc("shear")
    b.set_bc("shear")
    for s in range(8):
        H.move_linear(atoms, box, s, 1.0)
        L = H.build_list(atoms, box, 9.0)
        a.set_particles(L.tag, L.x, L.v, L.omega, L.radius)
        t = {k: DeviceArray(v) for k, v in
             dict(tag=L.tag.astype(np.int32), x=L.x, v=L.v, w=L.omega, r=L.radius).items()}
        b.set_particles_device(L.n, t["tag"].address, t["x"].address, t["v"].address,
                               t["w"].address, t["r"].address,
                               rmax_scale * float(L.radius.max()))
        a.step()
        b.step()
        Fa, Ta, Sa = a.hydro()
        Fb, Tb, Sb = b.hydro()  # synchronises b's stream before t goes out of scope
        assert np.array_equal(Fa, Fb) and np.array_equal(Ta, Tb) and np.array_equal(Sa, Sb)
    for x, y in zip(a.slots(), b.slots()):
        assert np.array_equal(x, y)
    assert np.array_equal(a.populations(), b.populations())
    a.close()
    b.close()
