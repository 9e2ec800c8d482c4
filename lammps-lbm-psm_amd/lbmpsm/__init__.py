"""lbmpsm — MI355X LBM-PSM fluid solver behind `fix lbm-psm` / `fix lbm-psm-bc`.

The compute path is the HIP library liblbmpsm.so (C ABI in include/lbmpsm.h);
this package is its Python handle plus the mock-LAMMPS harness used by the
tests and the benchmark.
"""
from . import _native
from ._native import BC_KEYWORDS, LbmPsmError
from .lattice import Lattice, config_from_fix, lattice_units_config

__all__ = ["Lattice", "config_from_fix", "lattice_units_config", "BC_KEYWORDS", "LbmPsmError",
           "_native"]
