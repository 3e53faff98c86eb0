"""MI355X (gfx950) Fock-build engine with the reference HF driver's surface.

The compute path is libgansu_fock.so (hand-written HIP kernels behind a C ABI,
include/gansu_fock.h); this package binds it with ctypes and provides the
workload geometry generators.
"""
from ._lib import GansuError, LIB_PATH, lib  # noqa: F401
from .engine import (FockEngine, Parameters, nuclear_repulsion_gradient, oneint, oneint_gradient, rhf,  # noqa: F401
                     rhf_gradient, rhf_orbitals, rohf, uhf, write_molden)
from .molecule import CONFIGS, Molecule, basis_path, config_molecule  # noqa: F401
