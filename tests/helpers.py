"""Test helpers: matching oracle / GPU states and layout conversions."""
import numpy as np

import oracle


def random_state(kind, lx, ly, lz, seed, amp=1e-3, umax=0.08, force=None):
    """Oracle Fields with a smooth random velocity field, rho ~ 1, populations =
    RR equilibrium + random non-equilibrium part, per-cell random force."""
    rng = np.random.default_rng(seed)
    g = oracle.Fields(kind, lx, ly, lz)
    n = g.n
    g.rho[:] = 1.0 + 0.01 * rng.standard_normal(n)
    g.ux[:] = umax * rng.uniform(-1, 1, n)
    g.uy[:] = umax * rng.uniform(-1, 1, n)
    g.uz[:] = umax * rng.uniform(-1, 1, n)
    g.init_equilibrium()
    _, w, _ = oracle.lattice(kind)
    g.f += amp * w[:, None] * rng.standard_normal((g.q, n))
    if force is None:
        g.fx[:] = 1e-4 * rng.standard_normal(n)
        g.fy[:] = 1e-4 * rng.standard_normal(n)
        g.fz[:] = 1e-4 * rng.standard_normal(n)
    else:
        g.set_force(force)
    g.taut[:] = 0.05 * rng.uniform(0, 1, n)
    return g


def upload_state(solver, g):
    from rrlbm import Field
    solver.upload(Field.F, g.f)
    solver.upload(Field.F_TEMP, g.ftmp)
    for fld, arr in ((Field.RHO, g.rho), (Field.UX, g.ux), (Field.UY, g.uy), (Field.UZ, g.uz),
                     (Field.FORCEX, g.fx), (Field.FORCEY, g.fy), (Field.FORCEZ, g.fz), (Field.TAUT, g.taut)):
        solver.upload(fld, arr)


def channel_state(kind, lx, ly, lz, u_tau, visc, seed, force):
    """Seeded channel initial state (the library's channel initialiser restated by the oracle)."""
    g = oracle.Fields(kind, lx, ly, lz)
    g.init_channel_fields(u_tau, visc, seed)
    g.init_equilibrium()
    g.set_force(force)
    return g


def maxabs(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))))


def fortran_to_cuda_order(f_fortran_aos, L, LZ):
    """Populations of the Fortran TG layout f(ip,i,j,k) (ip fastest, Fortran
    D3Q27 order) -> the library's [Q, N] SoA in the channel order with
    idx = (k*L + j)*L + i."""
    cF, _, _ = oracle.lattice(oracle.D3Q27_FORTRAN)
    cC, _, _ = oracle.lattice(oracle.D3Q27_CUDA)
    q = 27
    n = L * L * LZ
    aos = f_fortran_aos.reshape(n, q)  # cell index (i fastest) = i + L*(j + L*k) = our idx
    out = np.zeros((q, n))
    for ic in range(q):
        jf = int(np.where((cF == cC[ic]).all(axis=1))[0][0])
        out[ic] = aos[:, jf]
    return out


CAVITY_FIXTURE = "cavity_128x128x2_re1000_it5000_lines.txt"
# the reference Cavity solver's configuration (Cavity/para.f90:25-28, var_inc.f90:5,20):
# 128 x 128 x 2, lid speed 0.1 in +y on the x = LX face, Re 1000
CAVITY_N, CAVITY_LID, CAVITY_RE, CAVITY_STEPS = 128, 0.1, 1000.0, 5000


def cavity_tau():
    return 3.0 * (CAVITY_LID * CAVITY_N / CAVITY_RE) + 0.5


def cavity_fixture_errors(rho, ux, uy, golden_dir):
    """Largest |difference| of rho, ux, uy (reference layout, N = 128*128*2)
    from the reference Cavity solver's step-5000 fields on the fixture's lines."""
    import os
    a = np.loadtxt(os.path.join(golden_dir, CAVITY_FIXTURE))
    n = CAVITY_N
    axis, k, j = a[:, 0].astype(int), a[:, 1].astype(int), a[:, 2].astype(int)
    x = np.where(axis == 0, k, j)
    y = np.where(axis == 0, j, k)
    idx = y * n + x  # layer z = 0
    return tuple(float(np.abs(np.asarray(got)[idx] - a[:, c]).max()) for got, c in ((rho, 3), (ux, 4), (uy, 5)))
