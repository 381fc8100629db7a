"""Build recipe of the oracle (test infrastructure only).

  liboracle.so      -- the C restatement rrlbm_oracle.c (gcc, OpenMP, no FP contraction)
  _ref/             -- when /root/reference is present: the reference's own
                       Taylor-Green D3Q19 Fortran program compiled from its
                       source where it lies (amdflang, default REAL kind 4, as
                       the committed D3Q27 output was produced), run in _ref/
                       to produce D3Q19 output files.  Nothing from the
                       reference is copied into the repository.
  _ref/tg27/        -- (--tg27) the reference's D3Q27 Taylor-Green program,
                       built and run the same way; its output equals the
                       committed RRBMa0.01L100.txt column for column and its
                       first line times the reference's CPU path here.
  _ref/cavity/      -- the reference's lid-driven cavity solver (Cavity/*.f90,
                       128x128x2, D3Q27 RR, Re 1000, lid x = LX moving in +y)
                       compiled unmodified from its sources and run until its
                       first field dump (step 5000, Cavity/main.f90:37),
                       which tests/golden/make_fixtures.py turns into a
                       fixture.  Its declarations are type-consistent only
                       with 8-byte default REAL: collision.f90:9 and
                       streaming.f90:11 declare the REAL functions feqRR,
                       feqRR_4 `real*8`, and collision.f90 passes real*8
                       arguments to REAL dummies; so -fdefault-real-8.  -O0:
                       the -O2 build of this amdflang crashes at start-up.
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


def build_reference_tg27(run=True):
    """Compile Taylor_Green/D3Q27RR.f90 into _ref/tg27/ (amdflang -O2, default
    REAL kind 4) and (optionally) run it there: 20 000 steps of the 100 x 100 x 2
    Taylor-Green vortex, serial, a few minutes; it writes RRBMa0.01L100*.txt
    with its own execution time on the first line.  This build reproduces the
    committed Taylor_Green/RRBMa0.01L100.txt column for column, which also
    times the reference's own CPU path on this machine.

    Returns the output file, or None when the reference or amdflang is absent."""
    src = os.path.join(REFERENCE, "Taylor_Green", "D3Q27RR.f90")
    if not (os.path.exists(src) and os.path.exists(AMDFLANG)):
        return None
    work = os.path.join(REF_DIR, "tg27")
    os.makedirs(work, exist_ok=True)
    exe = os.path.join(work, "tg_d3q27")
    if not os.path.exists(exe):
        subprocess.run([AMDFLANG, "-O2", "-o", exe, src], check=True, cwd=work)
    out = os.path.join(work, "RRBMa0.01L100.txt")
    if run and not os.path.exists(out):
        subprocess.run([exe], check=True, cwd=work, stdout=subprocess.DEVNULL)
    return out


def same_midline(path_a, path_b):
    """Largest difference between two RRBMa0.01L100.txt midline tables."""
    import numpy as np

    def rows(p):
        out = []
        with open(p) as fh:
            for line in fh:
                t = line.split()
                if t and not t[0].startswith("#") and t[0] != "Execution":
                    out.append([float(v) for v in t])
        return np.array(out)

    a, b = rows(path_a), rows(path_b)
    return float(np.abs(a - b).max()) if a.shape == b.shape else float("inf")


CAVITY_SOURCES = ["var_inc.f90", "para.f90", "RR.f90", "collision.f90", "initial.f90", "streaming.f90", "main.f90"]
CAVITY_DUMP_STEP = 5000
CAVITY_CELLS = 128 * 128 * 2


def build_reference_cavity(run=True, timeout_s=5400):
    """Compile Cavity/*.f90 into _ref/cavity/ and (optionally) run it there
    until its step-5000 dump (32re1000/{rho,ux,uy,uz}5000.dat) is complete;
    the program itself runs 100 000 steps, so it is stopped then.

    Returns the dump directory, or None when the reference or amdflang is absent."""
    import time

    src_dir = os.path.join(REFERENCE, "Cavity")
    if not (os.path.isdir(src_dir) and os.path.exists(AMDFLANG)):
        return None
    work = os.path.join(REF_DIR, "cavity")
    os.makedirs(work, exist_ok=True)
    exe = os.path.join(work, "cavity")
    if not os.path.exists(exe):
        subprocess.run([AMDFLANG, "-O0", "-fdefault-real-8"] + [os.path.join(src_dir, f) for f in CAVITY_SOURCES]
                       + ["-o", exe], check=True, cwd=work)
    dump = os.path.join(work, "32re1000")
    names = [os.path.join(dump, "%s%d.dat" % (q, CAVITY_DUMP_STEP)) for q in ("rho", "ux", "uy", "uz")]
    whole = 8 * CAVITY_CELLS + 8  # one Fortran record: two 4-byte markers around the data

    def complete():
        return all(os.path.exists(n) and os.path.getsize(n) == whole for n in names)

    if not run or complete():
        return dump
    for n in names:
        if os.path.exists(n):
            os.remove(n)
    proc = subprocess.Popen([exe], cwd=work, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    t0 = time.time()
    try:
        while not complete():
            if proc.poll() is not None:
                raise RuntimeError("the reference Cavity program ended (status %s) before step %d"
                                   % (proc.returncode, CAVITY_DUMP_STEP))
            if time.time() - t0 > timeout_s:
                raise RuntimeError("the reference Cavity program did not reach step %d in %d s"
                                   % (CAVITY_DUMP_STEP, timeout_s))
            time.sleep(5)
        time.sleep(2)  # the last record is written; let the program close the file
    finally:
        if proc.poll() is None:
            proc.kill()
            proc.wait()
    return dump


if __name__ == "__main__":
    build_oracle()
    if "--ref" in sys.argv:
        print("reference build:", build_reference_tg19())
    if "--cavity" in sys.argv:
        print("reference cavity dump:", build_reference_cavity())
    if "--tg27" in sys.argv:
        out = build_reference_tg27()
        if out:
            with open(out) as fh:
                print(fh.readline().strip())
            print("largest difference from the committed output:",
                  same_midline(out, os.path.join(REFERENCE, "Taylor_Green", "RRBMa0.01L100.txt")))
