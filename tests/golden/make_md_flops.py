"""Writes md_flops.json: the FP64 work of one primitive quartet of each
four-centre class (orbital shells up to d) in the oracle's McMurchie-Davidson
routine (`classes`) and in its nuclear-derivative routine
(`gradient_classes`: the R table one order higher, Gamma(D) and the 12 centre
derivatives of every component), counted by running them on op-counting numbers
(oracle orc_md_flops: Boys, R recursion, E products, the Hermite contraction
sum, the prefactor, and the multiply-add that accumulates the primitive block
into the contracted one).  The reference-style Boys function is counted at
x = 1.0 (its table-Taylor branch from n = 3 on, the closed form with upward
recursion below).

Test infrastructure: run from the repository root in the build container
(`python tests/golden/make_md_flops.py`).  bench.py multiplies these counts by
the primitive quartets each class visits (gansu_fock_class_stats) for the
algorithmic FLOP rate of the direct build.
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_lib as O  # noqa: E402

X_BOYS = 1.0


def classes(lmax=2):
    pairs = [(la, lb) for la in range(lmax + 1) for lb in range(la + 1)]  # pair class id la(la+1)/2 + lb
    return [pairs[a] + pairs[b] for a in range(len(pairs)) for b in range(a + 1)]


def main():
    tab = O.boys_table()
    out, grad = {}, {}
    for c in classes():
        arith, special = O.md_flops(*c, x_boys=X_BOYS, table=tab)
        out["%d %d %d %d" % c] = [arith, special]
        arith, special = O.grad_md_flops(*c, x_boys=X_BOYS, table=tab)
        grad["%d %d %d %d" % c] = [arith, special]
    doc = dict(x_boys=X_BOYS, unit="FP64 operations per primitive quartet (arithmetic, special-function calls)",
               classes=out, gradient_classes=grad)
    with open(os.path.join(HERE, "md_flops.json"), "w") as f:
        json.dump(doc, f, indent=1, sort_keys=True)
        f.write("\n")


if __name__ == "__main__":
    main()
