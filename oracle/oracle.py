"""ctypes wrapper of liboracle.so -- the CPU restatement of the reference.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, as the oracle.  Arrays are numpy FP64 in the
channel solver's layout (populations [Q, N], idx = (iz*LY + iy)*LX + ix);
the Taylor-Green functions use the Fortran layout f(ip,i,j,k).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

D3Q27_CUDA, D3Q27_FORTRAN, D3Q19 = 0, 1, 2

_lib = None
_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.c_int


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            from build_oracle import build_oracle  # noqa: E402 (same directory)
            build_oracle(verbose=False)
        _lib = ctypes.CDLL(LIB_PATH)
        _lib.orc_uniform.restype = ctypes.c_double
        _lib.orc_uniform.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64]
    return _lib


def _p(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_D)


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def q_of(kind):
    return 19 if kind == D3Q19 else 27


def lattice(kind):
    q = q_of(kind)
    cx, cy, cz, opp = (np.zeros(q, dtype=np.int32) for _ in range(4))
    w = np.zeros(q)
    load().orc_lattice(kind, _ip(cx), _ip(cy), _ip(cz), _p(w), _ip(opp))
    return np.stack([cx, cy, cz], axis=1), w, opp


def hermite(kind):
    q = q_of(kind)
    out = np.zeros(16 * q)
    load().orc_hermite(kind, _p(out))
    return out.reshape(q, 16)


def hermite_high(kind):
    """4th-6th order tables H220 H202 H022 H211 H121 H112 H221 H212 H122 H222."""
    q = q_of(kind)
    out = np.zeros(10 * q)
    load().orc_hermite_high(kind, _p(out))
    return out.reshape(q, 10)


def set_threads(n):
    """Threads of the oracle's OpenMP loops."""
    load().orc_set_threads(ctypes.c_int(int(n)))


def get_threads():
    return int(load().orc_get_threads())


def set_rr_order(order):
    """3: the reference's RR operator; 6: with its (commented) 4th-6th order terms."""
    load().orc_set_rr_order(_I(order))


def get_rr_order():
    return load().orc_get_rr_order()


class rr_order:
    """Context manager: the oracle's RR routines at the given order inside."""

    def __init__(self, order):
        self.order = order

    def __enter__(self):
        self.prev = get_rr_order()
        set_rr_order(self.order)
        return self

    def __exit__(self, *exc):
        set_rr_order(self.prev)
        return False


def feq(kind, rho, u, v, w):
    out = np.zeros(q_of(kind))
    load().orc_feq(_I(kind), ctypes.c_double(rho), ctypes.c_double(u), ctypes.c_double(v), ctypes.c_double(w),
                   _p(out))
    return out


def rr_cell(kind, f, rho, u, v, w, F, tau_eff, mode):
    """One cell: mode 0 collision (collision.cu), mode 1 RRBC reconstruction (BC.cu)."""
    f = np.ascontiguousarray(f, dtype=np.float64)
    out = np.zeros(q_of(kind))
    d = ctypes.c_double
    load().orc_rr_cell(_I(kind), _p(f), d(rho), d(u), d(v), d(w), d(F[0]), d(F[1]), d(F[2]), d(tau_eff),
                       _I(mode), _p(out))
    return out


class Fields:
    """Reference-layout FP64 fields of one grid, stepped by the oracle."""

    def __init__(self, kind, lx, ly, lz):
        self.kind, self.lx, self.ly, self.lz = kind, lx, ly, lz
        self.q = q_of(kind)
        self.n = lx * ly * lz
        n = self.n
        self.f = np.zeros((self.q, n))
        self.ftmp = np.zeros((self.q, n))
        self.rho = np.ones(n)
        self.ux, self.uy, self.uz = np.zeros(n), np.zeros(n), np.zeros(n)
        self.fx, self.fy, self.fz = np.zeros(n), np.zeros(n), np.zeros(n)
        self.taut = np.zeros(n)

    def copy(self):
        g = Fields(self.kind, self.lx, self.ly, self.lz)
        for k in ("f", "ftmp", "rho", "ux", "uy", "uz", "fx", "fy", "fz", "taut"):
            setattr(g, k, getattr(self, k).copy())
        return g

    @property
    def dims(self):
        return (self.lx, self.ly, self.lz)

    # routines
    def forcing(self, force_in_x):
        load().orc_forcing(*self.dims, _p(self.fx), _p(self.fy), _p(self.fz), ctypes.c_double(force_in_x))

    def set_force(self, F):
        self.fx[:] = F[0]; self.fy[:] = F[1]; self.fz[:] = F[2]

    def forcing_gp(self, force_in_x, istep, phase):
        load().orc_forcing_gp(*self.dims, _p(self.fx), _p(self.fy), _p(self.fz), ctypes.c_double(force_in_x),
                              _I(istep), ctypes.c_double(phase))

    def init_arrays(self, ustar, ystar, force_in_x):
        d = ctypes.c_double
        load().orc_init_arrays(self.kind, *self.dims, _p(self.f), _p(self.rho), _p(self.ux), _p(self.uy),
                               _p(self.uz), _p(self.fx), _p(self.fy), _p(self.fz), d(ustar), d(ystar),
                               d(force_in_x))

    def init_equilibrium(self):
        load().orc_init_equilibrium(self.kind, *self.dims, _p(self.rho), _p(self.ux), _p(self.uy), _p(self.uz),
                                    _p(self.f))

    def init_channel_fields(self, ustar, visc, seed, amp_streak=0.1, amp_noise=0.01):
        d = ctypes.c_double
        load().orc_init_channel_fields(*self.dims, d(ustar), d(visc), ctypes.c_uint64(seed), d(amp_streak),
                                       d(amp_noise), _p(self.rho), _p(self.ux), _p(self.uy), _p(self.uz))

    def init_tg_fields(self, u0, three_d=False):
        load().orc_init_tg_fields(_I(int(three_d)), *self.dims, ctypes.c_double(u0), _p(self.rho), _p(self.ux),
                                  _p(self.uy), _p(self.uz))

    def les_wale(self, cw=0.5, guard=1):
        load().orc_les_wale(*self.dims, _p(self.ux), _p(self.uy), _p(self.uz), _p(self.taut), ctypes.c_double(cw),
                            _I(guard))

    def collision(self, tau, use_les=False):
        load().orc_collision_rr(self.kind, *self.dims, _p(self.f), _p(self.fx), _p(self.fy), _p(self.fz),
                                _p(self.rho), _p(self.ux), _p(self.uy), _p(self.uz),
                                _p(self.taut) if use_les else None, ctypes.c_double(tau))

    def streaming(self):
        load().orc_streaming(self.kind, *self.dims, _p(self.f), _p(self.ftmp))

    def bb(self, walls=(0, 1, 0), lid_face=-1, lid_u=(0.0, 0.0, 0.0)):
        lu = np.array(lid_u, dtype=np.float64)
        load().orc_bb(self.kind, *self.dims, _p(self.f), _p(self.ftmp), *[_I(w) for w in walls], _I(lid_face),
                      _p(lu))

    def swap(self):
        self.f, self.ftmp = self.ftmp, self.f

    def macrovar(self, wall_only=0, walls=(0, 0, 0)):
        load().orc_macrovar(self.kind, *self.dims, _p(self.f), _p(self.fx), _p(self.fy), _p(self.fz), _p(self.rho),
                            _p(self.ux), _p(self.uy), _p(self.uz), _I(wall_only), *[_I(w) for w in walls])

    def rrbc(self, walls=(0, 1, 0), lid_face=-1, lid_u=(0.0, 0.0, 0.0)):
        lu = np.array(lid_u, dtype=np.float64)
        load().orc_rrbc(self.kind, *self.dims, _p(self.f), _p(self.ftmp), _p(self.fx), _p(self.fy), _p(self.fz),
                        _p(self.rho), _p(self.ux), _p(self.uy), _p(self.uz), *[_I(w) for w in walls], _I(lid_face),
                        _p(lu))

    def fdbc(self, tau, use_les=False):
        """FDBC, BC.cu:129-309 (y walls; wall velocities extrapolated first)."""
        assert self.kind == D3Q27_CUDA
        load().orc_fdbc(*self.dims, _p(self.f), _p(self.fx), _p(self.fy), _p(self.fz), _p(self.ux), _p(self.uy),
                        _p(self.uz), _p(self.taut) if use_les else None, ctypes.c_double(tau))

    def step(self, nsteps, tau, use_les=False, cw=0.5, walls=(0, 1, 0), wall_rr=1, lid_face=-1,
             lid_u=(0.0, 0.0, 0.0)):
        """main.cu:50-84 order, nsteps times (wall_rr 2: FDBC in place of RRBC)."""
        lu = np.array(lid_u, dtype=np.float64)
        load().orc_step_channel(self.kind, *self.dims, _I(nsteps), _p(self.f), _p(self.ftmp), _p(self.fx),
                                _p(self.fy), _p(self.fz), _p(self.rho), _p(self.ux), _p(self.uy), _p(self.uz),
                                _p(self.taut), ctypes.c_double(tau), _I(int(use_les)), ctypes.c_double(cw), _I(1),
                                *[_I(w) for w in walls], _I(wall_rr), _I(lid_face), _p(lu))
        # (the C loop exchanges the contents of f and ftmp, so self.f holds the
        # streamed populations and self.ftmp the post-collision ones)

    def step_laminar(self, nsteps, tau, stale_fbar=True):
        """The Laminar forced-channel solver's step (Laminar/collision.f90:8-175,
        main.f90:78-92): walls on both x faces, D3Q27 only; stale_fbar keeps
        the reference's wall-loop Fbar (left from the sweep's last node)."""
        assert self.kind == D3Q27_CUDA
        load().orc_step_laminar(*self.dims, _I(nsteps), _p(self.f), _p(self.ftmp), _p(self.fx), _p(self.fy),
                                _p(self.fz), _p(self.rho), _p(self.ux), _p(self.uy), _p(self.uz),
                                ctypes.c_double(tau), _I(int(stale_fbar)))

    def collision_bgk(self, tau):
        """collision_BGK, BGK/collision.cu:3-33 (D3Q27, shifted populations)."""
        assert self.kind == D3Q27_CUDA
        load().orc_collision_bgk(*self.dims, _p(self.f), _p(self.fx), _p(self.fy), _p(self.fz), _p(self.rho),
                                 _p(self.ux), _p(self.uy), _p(self.uz), ctypes.c_double(tau))

    def macrovar_bgk(self):
        load().orc_macrovar_bgk(*self.dims, _p(self.f), _p(self.fx), _p(self.fy), _p(self.fz), _p(self.rho),
                                _p(self.ux), _p(self.uy), _p(self.uz))

    def step_bgk(self, nsteps, tau, walls=(0, 1, 0)):
        """BGK/main.cu:49-70 order, nsteps times."""
        assert self.kind == D3Q27_CUDA
        load().orc_step_bgk(*self.dims, _I(nsteps), _p(self.f), _p(self.ftmp), _p(self.fx), _p(self.fy),
                            _p(self.fz), _p(self.rho), _p(self.ux), _p(self.uy), _p(self.uz), ctypes.c_double(tau),
                            *[_I(w) for w in walls])

    def diag(self):
        out = np.zeros(9)
        load().orc_diag(ctypes.c_long(self.n), _p(self.rho), _p(self.ux), _p(self.uy), _p(self.uz), _p(out))
        return out

    def profiles(self, with_taut=False):
        out = np.zeros(10 * self.ly)
        load().orc_profiles(*self.dims, _p(self.ux), _p(self.uy), _p(self.uz), _p(self.taut) if with_taut else None,
                            _p(out))
        return out.reshape(self.ly, 10)


def laminar_parameters(nx=60):
    """Laminar/para.f90:32-38: Rstar = 50, ustar = 0.05, visc = ustar nx / Rstar,
    tau = 3 visc + 1/2, force_in_y = 8 visc ustar / nx^2 (FP64 arithmetic)."""
    rstar, ustar = 50.0, 0.05
    visc = ustar * float(nx) / rstar
    tau = 3.0 * visc + 0.5
    force_y = 8.0 * visc * ustar / float(nx) / float(nx)
    return {"ustar": ustar, "visc": visc, "tau": tau, "force_y": force_y}


def laminar_start(nx=60, ny=20, nz=2):
    """Laminar/main.f90:34-55: u = 0, uniform force (0, force_y, 0), populations
    = rest equilibrium w_i (initial.f90:2-31), then macrovar (u = F/2)."""
    prm = laminar_parameters(nx)
    g = Fields(D3Q27_CUDA, nx, ny, nz)
    g.set_force((0.0, prm["force_y"], 0.0))
    _, w, _ = lattice(D3Q27_CUDA)
    g.f[:] = w[:, None]
    g.macrovar()
    return g, prm


def uniform(seed, stream, counter):
    return load().orc_uniform(seed, stream, counter)


def tg_reference(kind, L, LZ, nsteps, real4=True):
    """Fortran Taylor-Green run (Taylor_Green/D3Q27RR.f90 / D3Q19RR.f90):
    returns (f, rho, u, v, w) after nsteps, Fortran layout ([k][j][i] for fields)."""
    q = q_of(kind)
    n = L * L * LZ
    f = np.zeros(q * n)
    rho, u, v, w = np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(n)
    lib = load()
    lib.orc_tg_ref_init(_I(kind), _I(L), _I(LZ), _I(int(real4)), _p(f), _p(rho), _p(u), _p(v), _p(w))
    cs = 1.0 / np.sqrt(3.0)
    U0 = 1.0 * cs / 100.0
    visc = U0 * L / (200.0 / cs)
    tau = visc / (1.0 / 3.0) + 0.5
    if nsteps:
        lib.orc_tg_ref_run(_I(kind), _I(L), _I(LZ), _I(int(real4)), _I(nsteps), ctypes.c_double(tau), _p(f),
                           _p(rho), _p(u), _p(v), _p(w))
    return f, rho, u, v, w, tau
