"""Build recipe of the oracle (test infrastructure only).

  liboracle.so      -- the C restatement rrlbm_oracle.c (gcc, OpenMP, no FP contraction)
  _ref/             -- when /root/reference is present: the reference's own
                       Taylor-Green D3Q19 Fortran program compiled from its
                       source where it lies (amdflang, default REAL kind 4, as
                       the committed D3Q27 output was produced), run in _ref/
                       to produce D3Q19 output files.  Nothing from the
                       reference is copied into the repository.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")
REFERENCE = "/root/reference"
AMDFLANG = "/opt/rocm/bin/amdflang"


def build_oracle(verbose=True):
    src = os.path.join(HERE, "rrlbm_oracle.c")
    tmp = LIB + ".tmp"
    subprocess.run(["gcc", "-O2", "-fopenmp", "-ffp-contract=off", "-fPIC", "-shared", "-o", tmp, src, "-lm"],
                   check=True)
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


def build_reference_tg19(run=True):
    """Compile Taylor_Green/D3Q19RR.f90 into _ref/ and (optionally) run it there.

    Returns the _ref directory, or None when the reference or amdflang is absent."""
    src = os.path.join(REFERENCE, "Taylor_Green", "D3Q19RR.f90")
    if not (os.path.exists(src) and os.path.exists(AMDFLANG)):
        return None
    os.makedirs(REF_DIR, exist_ok=True)
    exe = os.path.join(REF_DIR, "tg_d3q19")
    if not os.path.exists(exe):
        subprocess.run([AMDFLANG, "-O2", "-o", exe, src], check=True, cwd=REF_DIR)
    if run and not os.path.exists(os.path.join(REF_DIR, "D3Q19RR33BMa0.01L50.txt")):
        subprocess.run([exe], check=True, cwd=REF_DIR, stdout=subprocess.DEVNULL)
    return REF_DIR


if __name__ == "__main__":
    build_oracle()
    if "--ref" in sys.argv:
        print("reference build:", build_reference_tg19())
