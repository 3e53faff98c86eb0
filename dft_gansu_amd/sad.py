"""SAD cache files: one solved atom per element, appended to the .sad file
beside a basis file (the workflow of script/generate_sad_cache.py with
HF_main --export_sad_cache, hf.hpp:372-402).

Each element of the basis file is solved alone at the origin, with RHF when
its atomic number is even and ROHF when it is odd, and its per-spin densities
are appended as a "<symbol> <M>" header followed by the M x M alpha and beta
rows ("%.10e"), the layout the SAD guess reads.
"""
import os
import re
import tempfile

import numpy as np

from ._lib import check, lib
from .engine import Parameters
from .molecule import Molecule

_ELEMENT_LINE = re.compile(r"^([A-Z][a-z]?)\s+0\s*$")
_SYMBOLS = ("H He Li Be B C N O F Ne Na Mg Al Si P S Cl Ar K Ca Sc Ti V Cr Mn Fe Co Ni Cu Zn Ga Ge As Se Br Kr "
            "Rb Sr Y Zr Nb Mo Tc Ru Rh Pd Ag Cd In Sn Sb Te I Xe Cs Ba La Ce Pr Nd Pm Sm Eu Gd Tb Dy Ho Er Tm Yb "
            "Lu Hf Ta W Re Os Ir Pt Au Hg Tl Pb Bi Po At Rn Fr Ra Ac Th Pa U Np Pu Am Cm Bk Cf Es Fm Md No Lr Rf "
            "Db Sg Bh Hs Mt Ds Rg Cn Nh Fl Mc Lv Ts Og").split()


def basis_elements(gbs_path):
    """Element symbols that open a block of the .gbs file ("<symbol> 0")."""
    out = []
    with open(gbs_path) as fh:
        for line in fh:
            m = _ELEMENT_LINE.match(line.strip())
            if m and m.group(1) not in out:
                out.append(m.group(1))
    return out


def cached_elements(sad_path):
    """Element symbols already present in a .sad file."""
    if not os.path.exists(sad_path):
        return []
    out = []
    with open(sad_path) as fh:
        lines = fh.read().split("\n")
    i = 0
    while i < len(lines):
        tok = lines[i].split()
        if len(tok) == 2 and tok[0][:1].isalpha():
            out.append(tok[0])
            i += 1 + 2 * int(tok[1])
        else:
            i += 1
    return out


def export_sad_cache(mol, params, path, aux=None):
    """Solve mol with params' method and write its per-spin densities to path
    (HF::generate_sad_cache layout).  Returns (Da, Db)."""
    check(lib().gansu_export_sad_cache(mol.handle, aux.handle if aux is not None else None, params.handle,
                                       os.fsencode(path)))
    with open(path) as fh:
        M = int(fh.readline())
        vals = np.array(fh.read().split(), dtype=np.float64)
    return vals[:M * M].reshape(M, M), vals[M * M:].reshape(M, M)


def generate_sad_file(gbs_path, elements=None, sad_path=None, normalization="reference", **params):
    """Append the SAD densities of the basis file's elements (or the given
    symbols) that the .sad file does not hold yet.  Returns the symbols added."""
    sad_path = sad_path or os.path.splitext(gbs_path)[0] + ".sad"
    have = set(cached_elements(sad_path))
    added = []
    for sym in (elements or basis_elements(gbs_path)):
        if sym in have:
            continue
        z = _SYMBOLS.index(sym) + 1
        method = "rhf" if z % 2 == 0 else "rohf"
        mol = Molecule.from_atoms([z], np.zeros((1, 3)), gbs_path, normalization=normalization)
        prm = Parameters(method=method, **params)
        with tempfile.TemporaryDirectory() as tmp:
            Da, Db = export_sad_cache(mol, prm, os.path.join(tmp, "temp_sad_cache.dat"))
        with open(sad_path, "a") as fh:
            fh.write(f"{sym} {Da.shape[0]}\n")
            np.savetxt(fh, Da, fmt="%.10e")
            np.savetxt(fh, Db, fmt="%.10e")
        added.append(sym)
    return added
