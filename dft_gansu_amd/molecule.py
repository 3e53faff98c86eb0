"""Molecules: the .xyz/.gbs input path (C ABI) and the seeded geometry generators
of the benchmark workloads.

Input semantics follow the reference: parseXYZ.cpp (Angstrom -> bohr with
1.8897259886, types.hpp:141-145), basis_set.cpp (.gbs), molecular.hpp:117-190
(primitive list sorted by shell type, ShellTypeInfo, electron counts).
"""
import ctypes
import math
import os

import numpy as np

from ._lib import PRIM_DTYPE, TYPE_DTYPE, check, lib, ptr

ANGSTROM_TO_BOHR = 1.8897259886
BASIS_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "basis")
DEFAULT_SEED = 20261203
SYMBOLS = {"H": 1, "C": 6, "N": 7, "O": 8}


def basis_path(name):
    """Path of a bundled .gbs file ("sto-3g", "cc-pvdz", "6-31g_st", "cc-pvdz-rifit")."""
    p = os.path.join(BASIS_DIR, name + ".gbs")
    if not os.path.exists(p):
        raise FileNotFoundError(p)
    return p


class Molecule:
    """Atoms + basis -> primitive shells (owned by the C library)."""

    def __init__(self, handle):
        self._h = handle
        sizes = np.zeros(7, dtype=np.int32)
        check(lib().gansu_molecule_sizes(self._h, ptr(sizes)))
        (self.n_prims, self.n_basis, self.n_types, self.n_atoms, self.n_electrons, self.n_alpha,
         self.n_beta) = [int(v) for v in sizes]
        self.prims = np.zeros(self.n_prims, PRIM_DTYPE)
        self.types = np.zeros(self.n_types, TYPE_DTYPE)
        self.norm = np.zeros(self.n_basis)
        self.Z = np.zeros(self.n_atoms, dtype=np.int32)
        self.xyz = np.zeros((self.n_atoms, 3))
        check(lib().gansu_molecule_get(self._h, ptr(self.prims), ptr(self.types), ptr(self.norm), ptr(self.Z),
                                       ptr(self.xyz)))

    @classmethod
    def from_xyz(cls, xyz_path, gbs_path, charge=0, beta_to_alpha=0, normalization="reference"):
        h = ctypes.c_void_p()
        check(lib().gansu_molecule_from_xyz(xyz_path.encode(), gbs_path.encode(), charge, beta_to_alpha,
                                            _norm_code(normalization), ctypes.byref(h)))
        return cls(h)._remember(gbs_path, charge, beta_to_alpha, normalization)

    @classmethod
    def from_atoms(cls, Z, xyz_bohr, gbs_path, charge=0, beta_to_alpha=0, normalization="reference"):
        Z = np.ascontiguousarray(Z, dtype=np.int32)
        xyz = np.ascontiguousarray(xyz_bohr, dtype=np.float64).reshape(-1)
        h = ctypes.c_void_p()
        check(lib().gansu_molecule_from_atoms(len(Z), ptr(Z), ptr(xyz), gbs_path.encode(), charge, beta_to_alpha,
                                              _norm_code(normalization), ctypes.byref(h)))
        return cls(h)._remember(gbs_path, charge, beta_to_alpha, normalization)

    def _remember(self, gbs_path, charge, beta_to_alpha, normalization):
        self.gbs_path, self.charge = gbs_path, charge
        self.beta_to_alpha, self.normalization = beta_to_alpha, normalization
        return self

    def with_coordinates(self, xyz_bohr):
        """The same atoms, basis and charge at new positions (bohr, n_atoms x 3)."""
        return Molecule.from_atoms(self.Z, xyz_bohr, self.gbs_path, self.charge, self.beta_to_alpha,
                                   self.normalization)

    @property
    def handle(self):
        return self._h

    def nuclear_repulsion(self):
        return lib().gansu_molecule_nuclear_repulsion(self._h)

    def sad_density(self):
        """(Da, Db): the atomic per-spin densities of the .sad file beside the
        basis file on the diagonal atom blocks (read_density_from_sad,
        utils.cpp:21-58; placement as rhf.hpp:551-584), M x M each."""
        Da = np.zeros((self.n_basis, self.n_basis))
        Db = np.zeros((self.n_basis, self.n_basis))
        check(lib().gansu_sad_density(self._h, ptr(Da), ptr(Db)))
        return Da, Db

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                lib().gansu_molecule_destroy(h)
            except Exception:
                pass
            self._h = None


def _norm_code(normalization):
    if normalization in ("reference", 0):
        return 0
    if normalization in ("unit", 1):
        return 1
    raise ValueError("normalization must be 'reference' or 'unit'")


# ----------------------------------------------------------- geometries ----
# All generators return (Z, xyz in bohr).  Geometry parameters in Angstrom.

def _rng(seed):
    return np.random.default_rng(DEFAULT_SEED if seed is None else seed)


def water(jitter=0.01, seed=None):
    """O at the origin, two H at 0.9572 A and 104.52 deg, seeded uniform jitter."""
    th = math.radians(104.52)
    r = 0.9572
    xyz = np.array([[0.0, 0.0, 0.0], [r, 0.0, 0.0], [r * math.cos(th), r * math.sin(th), 0.0]])
    if jitter:
        xyz = xyz + _rng(seed).uniform(-jitter, jitter, xyz.shape)
    return np.array([8, 1, 1]), xyz * ANGSTROM_TO_BOHR


def _tetra_dirs_ch2(n1, n2):
    u = -(n1 + n2)
    u /= np.linalg.norm(u)
    v = np.cross(n1, n2)
    v /= np.linalg.norm(v)
    a = math.radians(109.4712206) / 2
    return [u * math.cos(a) + v * math.sin(a), u * math.cos(a) - v * math.sin(a)]


def alkane(n_carbon=10, cc=1.54, ch=1.09, jitter=0.02, seed=None):
    """All-trans CnH(2n+2): planar zigzag carbon chain, tetrahedral angles."""
    th = math.radians(109.4712206)
    dx, dy = cc * math.sin(th / 2), cc * math.cos(th / 2) / 2
    C = np.array([[i * dx, dy if i % 2 else -dy, 0.0] for i in range(n_carbon)])
    H = []
    for i in range(n_carbon):
        nb = [C[j] - C[i] for j in (i - 1, i + 1) if 0 <= j < n_carbon]
        nb = [v / np.linalg.norm(v) for v in nb]
        if len(nb) == 2:
            for d in _tetra_dirs_ch2(nb[0], nb[1]):
                H.append(C[i] + ch * d)
        else:  # terminal CH3, staggered: one H in the chain plane, anti to the C-C bond's neighbour side
            n1 = nb[0]
            inplane = np.array([0.0, -1.0 if C[i][1] > 0 else 1.0, 0.0])
            inplane -= n1 * np.dot(inplane, n1)
            inplane /= np.linalg.norm(inplane)
            e2 = np.cross(n1, inplane)
            for k in range(3):
                phi = 2 * math.pi * k / 3
                d = n1 * math.cos(th) + math.sin(th) * (math.cos(phi) * inplane + math.sin(phi) * e2)
                H.append(C[i] + ch * d)
    xyz = np.vstack([C, np.array(H)])
    if jitter:
        xyz = xyz + _rng(seed).uniform(-jitter, jitter, xyz.shape)
    Z = np.array([6] * n_carbon + [1] * len(H))
    return Z, xyz * ANGSTROM_TO_BOHR


def coronene(cc=1.42, ch=1.09):
    """Planar coronene-like C24H12 cut from a hexagonal lattice around a ring centre."""
    a = cc
    # graphene lattice: atoms of the honeycomb around a hexagon centre at the origin
    pts = []
    a1 = np.array([1.5 * a, math.sqrt(3) / 2 * a])
    a2 = np.array([1.5 * a, -math.sqrt(3) / 2 * a])
    basis = [np.array([a, 0.0]), np.array([a / 2, math.sqrt(3) / 2 * a])]
    for i in range(-6, 7):
        for j in range(-6, 7):
            for b in basis:
                pts.append(i * a1 + j * a2 + b)
    pts = np.array(pts)
    r = np.linalg.norm(pts, axis=1)
    order = np.argsort(r, kind="stable")
    C = pts[order[:24]]
    H = []
    for i, c in enumerate(C):
        d = np.linalg.norm(C - c, axis=1)
        nb = [C[j] - c for j in range(len(C)) if j != i and d[j] < 1.1 * a]
        if len(nb) == 2:
            u = -(nb[0] + nb[1])
            u /= np.linalg.norm(u)
            H.append(c + ch * u)
    C3 = np.hstack([C, np.zeros((24, 1))])
    H3 = np.hstack([np.array(H), np.zeros((len(H), 1))])
    Z = np.array([6] * 24 + [1] * len(H))
    return Z, np.vstack([C3, H3]) * ANGSTROM_TO_BOHR


def _random_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def water_box(n=64, box=12.4, min_oo=2.7, seed=None):
    """n waters (config-0 geometry, no jitter) at random orientations, O positions
    rejection-sampled in a cubic box with a minimum O-O distance."""
    rng = _rng(seed)
    Zw, w = water(jitter=0.0)
    w = w / ANGSTROM_TO_BOHR
    O = []
    tries = 0
    while len(O) < n:
        tries += 1
        if tries > 1000000:
            raise RuntimeError("could not place waters")
        p = rng.uniform(0, box, 3)
        if all(np.linalg.norm(p - q) >= min_oo for q in O):
            O.append(p)
    Z, xyz = [], []
    for p in O:
        R = _random_rotation(rng)
        for k in range(3):
            Z.append(int(Zw[k]))
            xyz.append(p + R @ w[k])
    return np.array(Z), np.array(xyz) * ANGSTROM_TO_BOHR


# workload configurations (sizes quoted Cartesian, as the reference computes)
CONFIGS = {
    0: dict(name="H2O/STO-3G", geom=lambda: water(), basis="sto-3g", aux=None),
    1: dict(name="C10H22/cc-pVDZ", geom=lambda: alkane(10), basis="cc-pvdz", aux=None),
    2: dict(name="C24H12/cc-pVDZ+RIFIT", geom=lambda: coronene(), basis="cc-pvdz", aux="cc-pvdz-rifit"),
    3: dict(name="C10H22/6-31G*", geom=lambda: alkane(10), basis="6-31g_st", aux=None),
    4: dict(name="(H2O)64/cc-pVDZ+RIFIT", geom=lambda: water_box(), basis="cc-pvdz", aux="cc-pvdz-rifit"),
}


def config_molecule(cfg, normalization="reference"):
    c = CONFIGS[cfg]
    Z, xyz = c["geom"]()
    mol = Molecule.from_atoms(Z, xyz, basis_path(c["basis"]), normalization=normalization)
    aux = Molecule.from_atoms(Z, xyz, basis_path(c["aux"]), normalization=normalization) if c["aux"] else None
    return mol, aux
