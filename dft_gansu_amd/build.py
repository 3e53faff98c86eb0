"""Build libgansu_fock.so in-tree for gfx950 (hipcc via the csrc Makefile)."""
import os
import subprocess
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG_DIR, "csrc")


def build(jobs=8, force=True, verbose=False):
    env = dict(os.environ)
    if force:
        subprocess.check_call(["make", "-s", "clean"], cwd=CSRC, env=env)
    cmd = ["make", f"-j{jobs}"]
    r = subprocess.run(cmd, cwd=CSRC, env=env, capture_output=not verbose, text=True)
    if r.returncode != 0:
        sys.stderr.write((r.stdout or "") + (r.stderr or ""))
        raise RuntimeError("libgansu_fock.so build failed")
    return os.path.join(PKG_DIR, "libgansu_fock.so")


if __name__ == "__main__":
    print(build(verbose=True))
