"""ctypes binding of libgansu_fock.so (the C ABI declared in include/gansu_fock.h).

The library is required: importing a compute entry point without the built
library raises immediately; there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
# GANSU_FOCK_LIB points at another build of the same library (e.g. an experiment).
LIB_PATH = os.environ.get("GANSU_FOCK_LIB") or os.path.join(PKG_DIR, "libgansu_fock.so")

PRIM_DTYPE = np.dtype([("exponent", "f8"), ("coefficient", "f8"), ("x", "f8"), ("y", "f8"), ("z", "f8"),
                       ("shell_type", "i4"), ("_pad0", "i4"), ("basis_index", "u8"), ("atom_index", "i4"),
                       ("_pad1", "i4")])
TYPE_DTYPE = np.dtype([("count", "i4"), ("_pad", "i4"), ("start_index", "u8")])
assert PRIM_DTYPE.itemsize == 64 and TYPE_DTYPE.itemsize == 16

# error codes of gansu_fock.h
ERRORS = {1: "GANSU_E_ARG", 2: "GANSU_E_IO", 3: "GANSU_E_BASIS", 4: "GANSU_E_ELECTRONS", 5: "GANSU_E_PARAM",
          6: "GANSU_E_HIP", 7: "GANSU_E_STATE", 8: "GANSU_E_NUMERIC", 9: "GANSU_E_UNSUPPORTED"}

# every symbol the header declares
EXPORTS = [
    "gansu_last_error", "gansu_version", "gansu_molecule_from_xyz", "gansu_molecule_from_atoms",
    "gansu_molecule_sizes", "gansu_molecule_get", "gansu_molecule_nuclear_repulsion", "gansu_molecule_destroy",
    "gansu_sad_density", "gansu_export_sad_cache",
    "gansu_params_create", "gansu_params_load_recipe", "gansu_params_set", "gansu_params_get", "gansu_params_destroy",
    "gansu_fock_create", "gansu_fock_precompute", "gansu_fock_build_rhf", "gansu_fock_build_jk", "gansu_fock_build_rhf_occ",
    "gansu_fock_algorithm_name", "gansu_fock_schwarz", "gansu_fock_num_primitive_pairs", "gansu_fock_stats",
    "gansu_fock_class_stats",
    "gansu_fock_set_profiling", "gansu_fock_last_profile", "gansu_fock_destroy", "gansu_oneint", "gansu_rhf_run",
    "gansu_fock_gradient_rhf", "gansu_fock_gradient_rhf_occ", "gansu_oneint_gradient", "gansu_molecule_nuclear_repulsion_gradient",
    "gansu_rhf_gradient", "gansu_oneint_noalloc", "gansu_rhf_run_guess", "gansu_rhf_gradient_guess", "gansu_uhf_run_guess", "gansu_uhf_run", "gansu_rohf_run", "gansu_rhf_run_mo", "gansu_write_molden", "gansu_fock_build_uhf", "gansu_fock_build_rohf", "gansu_fock_update_geometry",
]


class GansuError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


_LIB = None


def lib():
    """Load the HIP library; raise if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`"
                              " (there is no CPU fallback)")
        # torch ships its own HIP runtime; when the library's (system) runtime
        # is loaded first, torch's later initialisation sees no device and
        # the library's calls then fail, so torch is loaded first when present
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        vp, cp, ip, dp, sz = ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.c_size_t
        L.gansu_last_error.restype = cp
        L.gansu_version.restype = cp
        L.gansu_molecule_from_xyz.argtypes = [cp, cp, ip, ip, ip, ctypes.POINTER(vp)]
        L.gansu_molecule_from_atoms.argtypes = [ip, vp, vp, cp, ip, ip, ip, ctypes.POINTER(vp)]
        L.gansu_molecule_sizes.argtypes = [vp, vp]
        L.gansu_molecule_get.argtypes = [vp, vp, vp, vp, vp, vp]
        L.gansu_molecule_nuclear_repulsion.argtypes = [vp]
        L.gansu_molecule_nuclear_repulsion.restype = ctypes.c_double
        L.gansu_molecule_destroy.argtypes = [vp]
        L.gansu_sad_density.argtypes = [vp, vp, vp]
        L.gansu_export_sad_cache.argtypes = [vp, vp, vp, ctypes.c_char_p]
        L.gansu_params_create.argtypes = [ctypes.POINTER(vp)]
        L.gansu_params_load_recipe.argtypes = [vp, cp]
        L.gansu_params_set.argtypes = [vp, cp, cp]
        L.gansu_params_get.argtypes = [vp, cp, cp, ip]
        L.gansu_params_destroy.argtypes = [vp]
        L.gansu_fock_create.argtypes = [vp, sz, vp, ip, vp, ip, vp, sz, vp, ip, vp, ip, cp, ctypes.POINTER(vp)]
        L.gansu_fock_precompute.argtypes = [vp]
        L.gansu_fock_build_rhf.argtypes = [vp, vp, vp, vp, vp]
        L.gansu_fock_build_jk.argtypes = [vp, vp, vp, vp, vp]
        L.gansu_fock_build_rhf_occ.argtypes = [vp, vp, vp, ip, vp, vp, vp]
        L.gansu_fock_algorithm_name.argtypes = [vp]
        L.gansu_fock_algorithm_name.restype = cp
        L.gansu_fock_schwarz.argtypes = [vp, vp, sz]
        L.gansu_fock_num_primitive_pairs.argtypes = [vp]
        L.gansu_fock_num_primitive_pairs.restype = sz
        L.gansu_fock_stats.argtypes = [vp, vp]
        L.gansu_fock_class_stats.argtypes = [vp, ctypes.c_int, vp, vp, vp, vp]
        L.gansu_fock_set_profiling.argtypes = [vp, ip]
        L.gansu_fock_last_profile.argtypes = [vp, vp, ip, vp]
        L.gansu_fock_destroy.argtypes = [vp]
        L.gansu_oneint.argtypes = [vp, vp, vp, vp, vp]
        L.gansu_oneint_noalloc.argtypes = [vp, vp, vp, vp, vp]
        L.gansu_rhf_run.argtypes = [vp, vp, vp, vp, vp, vp]
        L.gansu_uhf_run.argtypes = [vp, vp, vp, vp, vp, vp, vp]
        L.gansu_rohf_run.argtypes = [vp, vp, vp, vp, vp, vp, vp]
        L.gansu_rhf_run_mo.argtypes = [vp, vp, vp, vp, vp, vp, vp]
        L.gansu_rhf_run_guess.argtypes = [vp, vp, vp, vp, vp, ctypes.c_int, vp, vp, vp]
        L.gansu_rhf_gradient_guess.argtypes = [vp, vp, vp, vp, vp, ctypes.c_int, vp, vp, vp, vp]
        L.gansu_uhf_run_guess.argtypes = [vp, vp, vp, vp, vp, ctypes.c_int, vp, vp, vp, vp]
        L.gansu_write_molden.argtypes = [vp, vp, vp, vp, ip, ctypes.c_char_p]
        L.gansu_fock_gradient_rhf.argtypes = [vp, vp, ip, vp, vp]
        L.gansu_fock_gradient_rhf_occ.argtypes = [vp, vp, vp, ip, ip, vp, vp]
        L.gansu_oneint_gradient.argtypes = [vp, vp, vp, vp, vp]
        L.gansu_molecule_nuclear_repulsion_gradient.argtypes = [vp, vp]
        L.gansu_rhf_gradient.argtypes = [vp, vp, vp, vp, vp, vp]
        L.gansu_fock_build_uhf.argtypes = [vp, vp, vp, vp, vp, vp, vp]
        L.gansu_fock_build_rohf.argtypes = [vp, vp, vp, vp, vp, vp, ctypes.c_char_p, ip, ip, vp, vp, vp, vp]
        L.gansu_fock_update_geometry.argtypes = [vp, vp, ip]
        _LIB = L
        del dp
    return _LIB


def check(rc):
    if rc != 0:
        raise GansuError(rc, lib().gansu_last_error().decode())


def ptr(a):
    """Address of a numpy array or a torch tensor (device pointers for tensors)."""
    if a is None:
        return None
    if hasattr(a, "data_ptr"):
        return ctypes.c_void_p(a.data_ptr())
    return a.ctypes.data_as(ctypes.c_void_p)
