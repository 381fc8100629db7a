"""The five benchmark configurations of the hot path, as ready-made solvers.

C1  quasi-2D Taylor-Green, D3Q27, 100x100x4, u0 = 0.01 cs, Re = 100 (L = 100)
C2  lid-driven cavity 128^3, D3Q19 / D3Q27, lid u = 0.1 in x on the top (y) face, Re = 1000
C3  turbulent channel DNS, D3Q27, Re_tau = 180, 384x128x192, u_tau = 0.0045
C4  turbulent channel LES-WALE, D3Q27, Re_tau = 180, 192x64x96, Cw = 0.5
C5  batched 3D Taylor-Green ensemble (see ensemble_instances)

Channel physics (C3/C4): h = LY/2, nu = u_tau h / 180, tau = 3 nu + 1/2,
F_x = u_tau^2 / h (globals.cu:57-63 with LY = 2h), seeded log-law + streak +
noise start.  The reference channel runs 64^3 at the same Re_tau with the
same force definition (RR27V33_code_DNS_LES/globals.cu:57-63).
"""
import math

import numpy as np

from .solver import Batch, Solver, Storage

CS = 1.0 / math.sqrt(3.0)


def taylor_green_2d(storage=Storage.F64, lx=100, ly=100, lz=4, ma=0.01, re=100.0, **kw):
    u0 = ma * CS
    nu = u0 * lx / re
    s = Solver(lx=lx, ly=ly, lz=lz, npop=27, storage=int(storage), tau=3 * nu + 0.5, force=(0, 0, 0), wall_y=0,
               wall_rr=0, **kw)
    s.init_taylor_green_2d(u0)
    return s, {"u0": u0, "nu": nu, "k": 2 * math.pi / lx}


def cavity(storage=Storage.F32, n=128, npop=27, lid=0.1, re=1000.0, wall_rr=1, **kw):
    nu = lid * n / re
    s = Solver(lx=n, ly=n, lz=n, npop=npop, storage=int(storage), tau=3 * nu + 0.5, force=(0, 0, 0), wall_x=1,
               wall_y=1, wall_z=1, wall_rr=wall_rr, lid_face=3, lid_u=(lid, 0.0, 0.0), **kw)
    s.init_rest()
    return s, {"nu": nu, "lid": lid}


def channel(storage=Storage.F32, lx=384, ly=128, lz=192, re_tau=180.0, u_tau=0.0045, les=False, seed=2024,
            amp_streak=0.1, amp_noise=0.01, wall_rr=1, **kw):
    h = ly / 2.0
    nu = u_tau * h / re_tau
    s = Solver(lx=lx, ly=ly, lz=lz, npop=27, storage=int(storage), tau=3 * nu + 0.5, force=(u_tau * u_tau / h, 0, 0),
               wall_y=1, wall_rr=wall_rr, use_les=int(les), cw=0.5, **kw)
    s.init_channel(u_tau, nu, seed, amp_streak, amp_noise)
    return s, {"nu": nu, "u_tau": u_tau, "h": h, "ystar": nu / u_tau}


def channel_les(storage=Storage.F16, **kw):
    return channel(storage=storage, lx=192, ly=64, lz=96, les=True, **kw)


def channel_bgk(storage=Storage.F32, **kw):
    """The channel DNS grid under the BGK baseline operator (BGK/collision.cu:
    second-order equilibrium, plain bounce-back walls)."""
    return channel(storage=storage, wall_rr=0, collision_model=1, **kw)


def ensemble_instances(seed=5, count=64):
    """64 periodic 3D Taylor-Green boxes: N in {32,48,64,96,128,192}, Ma in
    U[0.01, 0.1] (u0 = Ma cs), Re in U[100, 1600], nu = u0 (N / 2 pi) / Re."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = int(rng.choice([32, 48, 64, 96, 128, 192]))
        ma = rng.uniform(0.01, 0.1)
        re = rng.uniform(100.0, 1600.0)
        u0 = ma * CS
        nu = u0 * (n / (2 * math.pi)) / re
        out.append((n, n, n, 3.0 * nu + 0.5, u0))
    return out


def ensemble(storage=Storage.F32, keep_reference=False, **kw):
    inst = ensemble_instances(**kw)
    return Batch(inst, storage=int(storage), keep_reference=keep_reference), inst


def plane_mean_u(solver):
    """Plane-averaged streamwise velocity <u>(y) of the reference fields."""
    return solver.output_profile()[:, 0]
