"""Writes c10h22_631gs_gradient.json: the FP64 host oracle's RHF energy and
analytic nuclear gradient of the C10H22/6-31G* workload (the all-trans alkane
generator with its default seed, reference normalisation, Schwarz 1e-12,
SCF converged to 1e-12 Eh).

Test infrastructure: run from the repository root in the build container
(`python tests/golden/make_c10h22_631gs_gradient.py`); the GPU test compares
the library's analytic gradient with the committed numbers.  About ten
minutes on 8 host threads.
"""
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import oracle_lib as O  # noqa: E402
from dft_gansu_amd import molecule as gm  # noqa: E402


def geometry_digest(Z, xyz):
    return hashlib.sha256(np.ascontiguousarray(Z, dtype=np.int64).tobytes() +
                          np.ascontiguousarray(xyz, dtype=np.float64).tobytes()).hexdigest()


def main():
    Z, xyz = gm.CONFIGS[3]["geom"]()
    t0 = time.time()
    nocc = int(np.sum(Z)) // 2
    prims, types, norm = O.build_molecule(gm.basis_path("6-31g_st"), Z, xyz, quirk=True)
    # direct J/K per iteration (the dense tensor would take 11 GB at M = 194)
    _, E, D, F, it = O.rhf(prims, types, norm, Z, xyz, nocc, thr=1e-12, conv=1e-12, maxiter=200, dense=False)
    W = 0.5 * D @ F @ D
    Q = O.schwarz(prims, types)
    g = (O.grad_1e(prims, norm, Z, xyz, D, W) + O.grad_2e(prims, types, norm, Q, D, len(Z), 1e-12)
         + O.grad_nuc(Z, xyz))
    out = {
        "workload": "C10H22/6-31G* RHF, reference normalisation, Schwarz 1e-12, conv 1e-12",
        "geometry_sha256": geometry_digest(Z, xyz),
        "n_atoms": int(len(Z)),
        "e_total": float(E),
        "scf_iterations": int(it),
        "gradient": [[float(v) for v in row] for row in g],
        "oracle_seconds": round(time.time() - t0, 1),
        "threads": O.lib().orc_num_threads(),
    }
    with open(os.path.join(HERE, "c10h22_631gs_gradient.json"), "w") as fh:
        json.dump(out, fh, indent=1)
    print("E =", E, "max|g| =", np.abs(g).max(), "in", out["oracle_seconds"], "s")


if __name__ == "__main__":
    main()
