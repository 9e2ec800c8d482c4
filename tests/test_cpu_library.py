"""CPU-side checks of the HIP library and the harness (no GPU needed).

* the C-ABI library loads and exports every entry point include/lbmpsm.h
  declares;
* lbmpsm_config_from_fix reproduces the reference unit conversion bit for
  bit (checked against the oracle's restatement) and rejects what
  `fix lbm-psm` rejects;
* the library's solid-fraction routine, including the lane split the map
  kernels use, returns the oracle's exact values;
* the mock-LAMMPS harness (ghost images, RSA) behaves as documented.
"""
import ctypes as C
import math
import os
import re

import numpy as np
import pytest

import lbmpsm
import lbmpsm_oracle as O
from lbmpsm import _native as N
from lbmpsm import harness as H

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_abi_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "lbmpsm.h")).read()
    names = sorted(set(re.findall(r"\b(lbmpsm_[a-z_]+)\s*\(", hdr)))
    assert len(names) >= 18
    lib = C.CDLL(N.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_probe_library_exports_and_rejects_bad_shapes():
    """The measurement library (copy ceiling, collide access pattern) loads
    and checks its arguments before touching a device."""
    from lbmpsm import probe
    lib = probe._lib()
    assert hasattr(lib, "lbmpsm_copy_probe_form") and hasattr(lib, "lbmpsm_pull_probe")
    ms = C.c_double()
    assert lib.lbmpsm_pull_probe(16, 16, 100, 8, 1, C.byref(ms)) == -1  # z not whole blocks
    assert lib.lbmpsm_pull_probe(16, 16, 64, 3, 1, C.byref(ms)) == -1   # element size
    assert lib.lbmpsm_copy_probe_form(6, 1 << 20, 1, C.byref(ms)) == -1  # no such form


def _cfg(*args, **kw):
    return lbmpsm.config_from_fix(*args, **kw)


@pytest.mark.parametrize("deck", [
    # (dim, Nlc, lc, rho, nu, Re, tau, Fext, boxlo, prd, periodic)
    (3, 11, 0.0005, 1000.0, 1e-4, 0.1, 0.7, (0, 0, 0), (0, 0, 0), (0.005,) * 3, (1, 0, 1)),
    (2, 11, 0.0005, 1000.0, 1e-4, 0.1, 0.7, (0, 0, 0), (0.0, -0.0125, -0.00025),
     (0.07, 0.025, 0.0005), (1, 1, 1)),
    (3, 6, 0.0125, 100.0, 0.005, 0.1, 0.65, (-981.0, 0, 0), (0, -0.15, -0.15),
     (1.0, 0.3, 0.3), (0, 0, 0)),
])
def test_config_from_fix_matches_reference_units(deck):
    dim, Nlc, lc, rho, nu, Re, tau, Fext, lo, prd, per = deck
    cfg = _cfg(dim, Nlc, lc, rho, nu, Re, tau, Fext, lo, prd, per)
    u = O.units(rho, nu, lc, Re, Nlc, tau, dim)
    assert cfg.dx == u.dx and cfg.Uc == u.Uc and cfg.u_lb == u.u_lb
    assert cfg.tc == u.tc and cfg.dt_d == u.dt_d
    assert cfg.force_factor == u.forceFactor and cfg.torque_factor == u.torqueFactor
    for d in range(3):
        assert cfg.F_lbm[d] == Fext[d] / u.volumeForceFactor
    # node counts, src/fix_LBM_PSM.cpp:142-154
    for d, nd in enumerate((cfg.nx, cfg.ny, cfg.nz)):
        if d < dim:
            expect = int(prd[d] / u.dx + 1.5) - (1 if per[d] else 0)
            assert nd == expect
    assert cfg.origin[2] == (0.0 if dim == 2 else lo[2])


def test_config_rejects_what_the_fix_rejects():
    base = (3, 11, 0.0005, 1000.0, 1e-4, 0.1, 0.7, (0, 0, 0), (0, 0, 0), (0.005,) * 3,
            (1, 1, 1))
    with pytest.raises(lbmpsm.LbmPsmError):  # tau <= 0.5
        _cfg(*base[:6], 0.5, *base[7:])
    with pytest.raises(lbmpsm.LbmPsmError):  # Ma > 0.3 (u_lb = Re nu_lb / (N-1))
        _cfg(3, 11, 0.0005, 1000.0, 1e-4, 100.0, 0.7, *base[7:])
    with pytest.raises(lbmpsm.LbmPsmError):  # box not a whole number of cells
        _cfg(*base[:9], (0.00501, 0.005, 0.005), base[10])


def test_create_without_gpu_reports_nodevice():
    from lbmpsm.devmem import device_count
    if device_count() > 0:
        pytest.skip("a GPU is present")
    cfg = lbmpsm.lattice_units_config(3, (8, 8, 8), (1, 1, 1), 0.7)
    with pytest.raises(lbmpsm.LbmPsmError) as e:
        lbmpsm.Lattice(cfg)
    assert e.value.code == N.ERR_NODEVICE


@pytest.fixture(scope="module")
def sfcheck():
    path = os.path.join(ROOT, "tests", "native", "libsf_check.so")
    if not os.path.exists(path):
        pytest.skip("tests/native not built")
    lib = C.CDLL(path)
    args = [C.c_int, C.c_double, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
            C.c_double]
    lib.sf_fast.restype = C.c_double
    lib.sf_fast.argtypes = args
    lib.sf_split.restype = C.c_double
    lib.sf_split.argtypes = args + [C.c_int]
    lib.sf_columns.restype = C.c_double
    lib.sf_columns.argtypes = args + [C.c_int]
    return lib


@pytest.mark.parametrize("dim", [2, 3])
def test_library_solid_fraction_equals_oracle(sfcheck, dim):
    o = O.Oracle(dim, (30, 30, 30), (1, 1, 1), 0.8)
    hs = math.sqrt(3.1) / 2.0 if dim == 3 else math.sqrt(2.1) / 2.0
    rng = np.random.default_rng(7 + dim)
    for trial in range(3000):
        r = rng.uniform(0.3, 12.0)
        d = rng.standard_normal(3)
        if dim == 2:
            d[2] = 0.0
        d = d / np.linalg.norm(d) * (r + rng.uniform(-1.2 * hs, 1.2 * hs))
        node = np.array([15, 15, 15 if dim == 3 else 0])
        xl = node - d
        if dim == 2:
            xl[2] = 1.0  # the 2D quirk: z_loc = 1 while the node sits at k = 0
        # also exercise exact-grid positions (ties with sub-point offsets)
        if trial % 7 == 0:
            xl = np.round(xl * 20) / 20
        args = (int(node[0]), int(node[1]), int(node[2]), xl[0], xl[1], xl[2], r)
        expect = o.solid_fraction(*args)
        assert sfcheck.sf_fast(dim, hs, *args) == expect
        for groups in (1, 8, 16, 64):
            assert sfcheck.sf_split(dim, hs, *args, groups) == expect
        for axis in (-1, 0, 1, 2):
            assert sfcheck.sf_columns(dim, hs, *args, axis) == expect


def test_line_form_counts_whole_zones(sfcheck):
    """The 3D map kernel's line form (one interval per z line of sub-points)
    gives every counted cell of whole sphere zones the per-cell count, which
    test_library_solid_fraction_equals_oracle ties to the oracle."""
    sfcheck.sf_lines_plane.restype = C.c_int
    sfcheck.sf_lines_plane.argtypes = [C.c_double, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                       C.c_double, C.c_double, C.c_double, C.c_double]
    hs = math.sqrt(3.1) / 2.0
    rng = np.random.default_rng(31)
    planes = 0
    for trial in range(60):
        r = rng.uniform(0.6, 11.0)
        c = rng.uniform(20.0, 21.0, 3)
        if trial % 5 == 0:
            c = np.round(c * 4) / 4  # centres on sub-grid-aligned values
        lo = (c - r).astype(int) - 2
        hi = (c + r).astype(int) + 2
        for i in range(lo[0], hi[0]):
            bad = sfcheck.sf_lines_plane(hs, int(i), int(lo[1]), int(lo[2]), int(hi[1] - lo[1]),
                                         int(hi[2] - lo[2]), c[0], c[1], c[2], r)
            assert bad == 0, (trial, i)
            planes += 1
    assert planes > 600


def test_line_ends_hold_with_an_inexact_square_root():
    """The line form predicts both ends of a line from a float square root
    and skips the exact tests of an end that lies clear of a sub-point by
    more than its error bound.  With that square root off by up to 4 ulp
    (the GPU's is within 1-2), every counted cell still gets the per-cell
    count: zones at random, with centres on a quarter grid, and with centres
    and radii on the sub-point grid (crossings exactly on sub-points)."""
    path = os.path.join(ROOT, "tests", "native", "libsf_check_ulp.so")
    if not os.path.exists(path):
        pytest.skip("tests/native not built")
    lib = C.CDLL(path)
    lib.sf_lines_plane.restype = C.c_int
    lib.sf_lines_plane.argtypes = [C.c_double, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                   C.c_double, C.c_double, C.c_double, C.c_double]
    lib.sf_set_sqrt_ulps.argtypes = [C.c_int]
    hs = math.sqrt(3.1) / 2.0
    rng = np.random.default_rng(57)
    cases = []
    for trial in range(24):
        r = rng.uniform(0.6, 11.0)
        c = rng.uniform(20.0, 21.0, 3)
        if trial % 3 == 1:
            c = np.round(c * 4) / 4
        elif trial % 3 == 2:
            c = np.round(c * 20) / 20
            r = np.round(r * 20) / 20
        cases.append((c, r))
    planes = 0
    try:
        for ulps in (-4, -2, -1, 1, 2, 4):
            lib.sf_set_sqrt_ulps(ulps)
            for c, r in cases:
                lo = (c - r).astype(int) - 2
                hi = (c + r).astype(int) + 2
                for i in range(lo[0], hi[0]):
                    bad = lib.sf_lines_plane(hs, int(i), int(lo[1]), int(lo[2]),
                                             int(hi[1] - lo[1]), int(hi[2] - lo[2]),
                                             c[0], c[1], c[2], r)
                    assert bad == 0, (ulps, c, r, i)
                    planes += 1
    finally:
        lib.sf_set_sqrt_ulps(0)
    assert planes > 1000


def test_ghost_images_corner():
    box = H.Box(lo=np.zeros(3), prd=np.array([20.0, 20.0, 20.0]), periodic=(1, 1, 1))
    atoms = H.Atoms(tag=np.array([5], np.int32), x=np.array([[1.0, 19.0, 10.0]]),
                    v=np.zeros((1, 3)), omega=np.zeros((1, 3)), radius=np.array([2.0]))
    lst = H.build_list(atoms, box, cutoff=3.0)
    # x image (+20), y image (-20) of both, no z image: 4 entries
    assert lst.n == 4
    assert np.all(lst.tag == 5) and np.all(lst.owner == 0)
    got = {tuple(p) for p in lst.x}
    assert got == {(1.0, 19.0, 10.0), (21.0, 19.0, 10.0), (1.0, -1.0, 10.0), (21.0, -1.0, 10.0)}


def test_ghosts_skip_wall_axes_and_remap():
    box = H.Box(lo=np.zeros(3), prd=np.array([10.0, 9.0, 10.0]), periodic=(1, 0, 1))
    atoms = H.Atoms(tag=np.array([1], np.int32), x=np.array([[5.0, 0.5, 5.0]]),
                    v=np.array([[1.0, 0.0, -1.0]]), omega=np.zeros((1, 3)),
                    radius=np.array([1.0]))
    assert H.build_list(atoms, box, 2.0).n == 1
    H.move_linear(atoms, box, 7, 1.0)
    assert np.allclose(atoms.x, [[2.0, 0.5, 8.0]])


def test_rsa_no_overlap_and_seeded():
    box = H.Box(lo=np.zeros(3), prd=np.array([64.0, 63.0, 64.0]), periodic=(1, 0, 1))
    radii = np.full(40, 5.0)
    p1 = H.rsa_spheres(box, radii, 11, gap=0.5)
    p2 = H.rsa_spheres(box, radii, 11, gap=0.5)
    assert np.array_equal(p1, p2)
    d = p1[:, None, :] - p1[None, :, :]
    for a in (0, 2):
        d[..., a] -= 64.0 * np.round(d[..., a] / 64.0)
    dist = np.sqrt((d ** 2).sum(-1)) + np.eye(40) * 1e9
    assert dist.min() >= 10.5 - 1e-12
    assert np.all(p1[:, 1] >= 5.5) and np.all(p1[:, 1] <= 63.0 - 5.5)
