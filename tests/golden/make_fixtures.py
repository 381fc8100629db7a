"""Regenerate the golden fixtures in tests/golden/ (run in the build container).

Fixtures (plain text, data only):
  tg_d3q27_L100_it20000_fortran_midline.txt
      columns i, v, p, dvdx, dvdy of the reference's committed Taylor-Green
      output Taylor_Green/RRBMa0.01L100.txt (D3Q27RR.f90, 100x100x2, 20 000
      steps; midline = mean of rows j = L/2, L/2+1).
  tg_d3q19_L50_it10000_fortran_midline.txt
      the same columns from the reference's Taylor_Green/D3Q19RR.f90 compiled
      from source and run by oracle/build_oracle.py (oracle/_ref/, first case
      L = 50, 10 000 steps).
  tg2d_c1_ke_oracle.txt
      kinetic energy of the 100x100x4 Taylor-Green configuration under the
      oracle's reference-order step (steps 0, 500, ..., 2000).
  tg_d3q27_L100_it20000_oracle_fp64_midline.txt
      i, v, p of the oracle's pure-FP64 run of the same Taylor-Green case
      (for the FP64 GPU comparison; the Fortran output carries REAL*4
      intermediates).
  cavity_ghia_re1000.txt, cavity_ghia_re400.txt
      the Ghia et al. centre-line velocities the reference's Cavity solver
      ships for comparison (Cavity/dataref/Re1000.dat, RE400.dat).
  re180_spectral_uplus.txt
      y/h, y+, U+ of the spectral Re_tau = 180 channel profile the reference
      compares with (RR27V33_code_DNS_LES/Re180.prof).
  laminar_60x20x2_uy_profiles.txt
      the reference's committed FP64 fields of the Laminar forced channel
      (Laminar/output/u{x,y,z}{5000,10000,60000}.dat, Fortran unformatted,
      60x20x2, x fastest): the flow is uniform in y and z, so the file keeps
      the x profile of u_y at each step (all digits) and, in the header, the
      largest deviation of any node from its profile and the largest |u_x|,
      |u_z|, so a test can check the whole field.
  cavity_128x128x2_re1000_it5000_lines.txt
      the reference's lid-driven cavity solver (Cavity/*.f90, D3Q27 RR,
      128x128x2, walls on the x and y faces, lid x = LX moving in +y at 0.1,
      Re 1000) compiled from source and run by oracle/build_oracle.py to its
      first field dump (step 5000, Fortran unformatted rho, ux, uy, uz): rho,
      ux, uy along the lines x = 0, 1, 63, 64, 126, 127 and y = 0, 1, 63, 64,
      126, 127 of the z = 0 layer (all digits; 0-based indices, x fastest in
      the reference's arrays), with the largest |uz| and the largest
      difference between the two z layers in the header.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "oracle"))

REF_TG27 = "/root/reference/Taylor_Green/RRBMa0.01L100.txt"
COLS = "i v p dvdx dvdy"


def fortran_midline(src, dst, note):
    a = np.loadtxt(src, skiprows=2)
    out = np.column_stack([a[:, 0], a[:, 1], a[:, 3], a[:, 8], a[:, 9]])
    np.savetxt(dst, out, fmt=["%d", "%.10f", "%.10f", "%.10f", "%.10f"], header=note + "\ncolumns: " + COLS)


def oracle_fp64_midline(dst):
    import oracle
    L, LZ, it = 100, 2, 20000
    f, rho, u, v, w, tau = oracle.tg_reference(oracle.D3Q27_FORTRAN, L, LZ, it, real4=False)
    V = v.reshape(LZ, L, L)
    R = rho.reshape(LZ, L, L)
    vm = 0.5 * (V[0, L // 2 - 1, :] + V[0, L // 2, :])
    pm = 0.5 * ((R[0, L // 2 - 1, :] - 1.0) / 3.0 + (R[0, L // 2, :] - 1.0) / 3.0)
    np.savetxt(dst, np.column_stack([np.arange(1, L + 1), vm, pm]), fmt=["%d", "%.17e", "%.17e"],
               header="oracle FP64 Taylor-Green D3Q27 100x100x2, 20000 steps (D3Q27RR.f90 restated)\ncolumns: i v p")


def tg2d_c1_ke(dst):
    """Kinetic energy of the 100x100x4 Taylor-Green configuration (u0 = 0.01 cs,
    Re = 100) under the oracle's reference-order step, at steps 0, 500, ..., 2000."""
    import oracle
    lx, ly, lz = 100, 100, 4
    u0 = 0.01 / np.sqrt(3.0)
    nu = u0 * lx / 100.0
    tau = 0.5 + 3.0 * nu
    g = oracle.Fields(oracle.D3Q27_CUDA, lx, ly, lz)
    g.init_tg_fields(u0)
    g.init_equilibrium()
    rows = [(0, g.diag()[8])]
    for k in range(4):
        g.step(500, tau, walls=(0, 0, 0), wall_rr=0)
        rows.append((500 * (k + 1), g.diag()[8]))
    np.savetxt(dst, np.array(rows), fmt=["%d", "%.17e"],
               header="oracle FP64 kinetic energy, Taylor-Green 100x100x4, u0 = 0.01 cs, Re = 100\ncolumns: step KE")


REF_CAVITY = "/root/reference/Cavity/dataref"


def cavity_ghia(dst_dir):
    """The Ghia et al. centre-line data the reference's Cavity solver ships
    (Cavity/dataref/Re1000.dat, RE400.dat): columns y, u(x = 1/2, y) / U and
    x, v(x, y = 1/2) / U, lid at y = 1 moving in +x."""
    for src, re in (("Re1000.dat", 1000), ("RE400.dat", 400)):
        a = np.loadtxt(os.path.join(REF_CAVITY, src))
        np.savetxt(os.path.join(dst_dir, "cavity_ghia_re%d.txt" % re), a, fmt="%.5f",
                   header="Ghia et al. lid-driven cavity, Re = %d (Cavity/dataref/%s)\ncolumns: y u x v" % (re, src))


REF_RE180 = "/root/reference/RR27V33_code_DNS_LES/Re180.prof"


def re180_profile(dst):
    """Columns y/h, y+, U+ of the spectral Re_tau = 180 channel profile the
    reference compares its runs with (RR27V33_code_DNS_LES/Re180.prof,
    turbulent.ipynb)."""
    a = np.loadtxt(REF_RE180)
    np.savetxt(dst, a[:, :3], fmt="%.7e",
               header="spectral channel Re_tau = 180 mean profile (RR27V33_code_DNS_LES/Re180.prof)\ncolumns: y/h y+ U+")


REF_LAMINAR = "/root/reference/Laminar/output"


def read_fortran_record(path, count):
    """One Fortran unformatted sequential record of `count` FP64 values."""
    raw = open(path, "rb").read()
    n0 = int(np.frombuffer(raw[:4], dtype="<i4")[0])
    n1 = int(np.frombuffer(raw[4 + n0:8 + n0], dtype="<i4")[0])
    assert n0 == n1 == 8 * count and len(raw) == n0 + 8, (path, n0, n1, len(raw))
    return np.frombuffer(raw[4:4 + n0], dtype="<f8").copy()


def laminar_profiles(dst, nx=60, ny=20, nz=2):
    steps = (5000, 10000, 60000)
    prof, notes = [], []
    for it in steps:
        f = {c: read_fortran_record(os.path.join(REF_LAMINAR, "u%s%d.dat" % (c, it)), nx * ny * nz).reshape(nz, ny, nx)
             for c in "xyz"}
        p = f["y"][0, 0, :]
        dev = float(np.abs(f["y"] - p[None, None, :]).max())
        notes.append("step %d: max |uy - profile| = %.3e, max |ux| = %.3e, max |uz| = %.3e"
                     % (it, dev, float(np.abs(f["x"]).max()), float(np.abs(f["z"]).max())))
        prof.append(p)
    hdr = ("Laminar forced channel 60x20x2 (Laminar/), reference FP64 output u_y(x) at steps 5000, 10000, 60000\n"
           + "\n".join(notes) + "\ncolumns: ix uy_5000 uy_10000 uy_60000")
    np.savetxt(dst, np.column_stack([np.arange(nx)] + prof), fmt=["%d", "%.17e", "%.17e", "%.17e"], header=hdr)


CAVITY_LINES = (0, 1, 63, 64, 126, 127)


def cavity_reference_lines(dst, n=128, nz=2, step=5000):
    import build_oracle
    dump = build_oracle.build_reference_cavity(run=True)
    f = {q: read_fortran_record(os.path.join(dump, "%s%d.dat" % (q, step)), n * n * nz).reshape(nz, n, n)
         for q in ("rho", "ux", "uy", "uz")}
    layers = max(float(np.abs(f[q][1] - f[q][0]).max()) for q in f)
    rows = []
    for axis in (0, 1):  # 0: the line x = k along y; 1: the line y = k along x
        for k in CAVITY_LINES:
            for j in range(n):
                x, y = (k, j) if axis == 0 else (j, k)
                rows.append((axis, k, j, f["rho"][0, y, x], f["ux"][0, y, x], f["uy"][0, y, x]))
    hdr = ("reference Cavity solver (Cavity/*.f90 built from source, amdflang -O0 -fdefault-real-8), 128x128x2, "
           "D3Q27 RR, Re 1000, lid x = LX moving in +y at 0.1, fields after %d steps, layer z = 0\n"
           "max |uz| = %.3e, max difference between the z layers = %.3e\n"
           "columns: axis (0: line x = k, running y; 1: line y = k, running x) k j rho ux uy"
           % (step, float(np.abs(f["uz"]).max()), layers))
    np.savetxt(dst, np.array(rows), fmt=["%d", "%d", "%d", "%.17e", "%.17e", "%.17e"], header=hdr)


def main():
    fortran_midline(REF_TG27, os.path.join(HERE, "tg_d3q27_L100_it20000_fortran_midline.txt"),
                    "from Taylor_Green/RRBMa0.01L100.txt (reference output)")
    ref19 = os.path.join(REPO, "oracle", "_ref", "D3Q19RR33BMa0.01L50.txt")
    if not os.path.exists(ref19):
        import build_oracle
        build_oracle.build_reference_tg19(run=True)
    fortran_midline(ref19, os.path.join(HERE, "tg_d3q19_L50_it10000_fortran_midline.txt"),
                    "from Taylor_Green/D3Q19RR.f90 built from source (amdflang -O2, default REAL kind 4)")
    oracle_fp64_midline(os.path.join(HERE, "tg_d3q27_L100_it20000_oracle_fp64_midline.txt"))
    tg2d_c1_ke(os.path.join(HERE, "tg2d_c1_ke_oracle.txt"))
    laminar_profiles(os.path.join(HERE, "laminar_60x20x2_uy_profiles.txt"))
    cavity_ghia(HERE)
    cavity_reference_lines(os.path.join(HERE, "cavity_128x128x2_re1000_it5000_lines.txt"))
    re180_profile(os.path.join(HERE, "re180_spectral_uplus.txt"))


if __name__ == "__main__":
    main()
