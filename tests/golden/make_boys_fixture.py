"""Writes tests/golden/boys_ref_samples.json from the reference's own Boys grid.

The reference ships F_n(i/32), n = 0..29, i = 0..1023, as a C array of decimal
literals (src/boys/boys_30720.h, index n*1024 + i; hf.cu:58, :74-77).  This
script reads that file as text and records its digest plus a deterministic
sample of (n, i, value) entries.  Run it where the reference is mounted:
    python tests/golden/make_boys_fixture.py /root/reference
"""
import hashlib
import json
import os
import re
import sys


def main(ref_root):
    path = os.path.join(ref_root, "src", "boys", "boys_30720.h")
    text = open(path).read()
    decl = text[text.index("h_boys_grid"):]
    body = decl[decl.index("{") + 1:decl.index("}")]
    vals = [float(tok) for tok in re.findall(r"[-+]?[0-9]+(?:\.[0-9]*)?(?:[eE][-+]?[0-9]+)?", body)]
    assert len(vals) == 30720, len(vals)
    digest = hashlib.sha256(json.dumps(vals).encode()).hexdigest()
    samples = []
    for k in range(0, 30720, 53):
        samples.append([k // 1024, k % 1024, vals[k]])
    for n in range(30):  # first and last grid point of every order
        samples.append([n, 0, vals[n * 1024]])
        samples.append([n, 1023, vals[n * 1024 + 1023]])
    out = dict(source="src/boys/boys_30720.h", layout="n*1024+i, x_i=i/32", count=len(vals), sha256_json=digest,
               samples=samples)
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "boys_ref_samples.json")
    with open(dst, "w") as f:
        json.dump(out, f, indent=0)
    print(dst, len(samples))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
