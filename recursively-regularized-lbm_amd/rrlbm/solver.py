"""Host-side mirror of the channel solver's interface over the C ABI.

The reference drives its kernels from main() (RR27V33_code_DNS_LES/main.cu:31-117)
with the routine names collision_RR, streaming, BB, macrovar, RRBC,
macrovar_BN, LES_WALE, forcing, forcing_GP, init_arrays, diag_flow and
output_profile.  Solver keeps those names as methods, each a thin call into
librrlbm.so on the GPU; fields move as numpy FP64 arrays in the reference
layout (populations [NPOP, N], idx = (iz*LY + iy)*LX + ix).
"""
import ctypes
import enum

import numpy as np

from ._lib import Diag, Instance, Params, RRLBMError, Tuning, check, check_batch, load


class Storage(enum.IntEnum):
    F64 = 0
    F32 = 1
    F16 = 2


class Field(enum.IntEnum):
    F = 0
    F_TEMP = 1
    RHO = 2
    UX = 3
    UY = 4
    UZ = 5
    FORCEX = 6
    FORCEY = 7
    FORCEZ = 8
    TAUT = 9


def default_params(**overrides):
    """Params with the defaults of parameters.h / globals.cu, then overrides
    (force and lid_u accept 3-sequences)."""
    p = Params()
    check(load().rrlbm_params_default(ctypes.byref(p)))
    for k, v in overrides.items():
        if k in ("force", "lid_u"):
            arr = getattr(p, k)
            for i in range(3):
                arr[i] = float(v[i])
        elif hasattr(p, k):
            setattr(p, k, v)
        else:
            raise KeyError("unknown parameter %r" % k)
    return p


def lattice(npop):
    """(c[Q,3], w[Q], opp[Q]) of the library's velocity ordering."""
    c = (ctypes.c_int32 * (3 * npop))()
    w = (ctypes.c_double * npop)()
    o = (ctypes.c_int32 * npop)()
    check(load().rrlbm_lattice(npop, c, w, o))
    return (np.array(c, dtype=np.int32).reshape(npop, 3), np.array(w), np.array(o, dtype=np.int32))


def device_count():
    n = ctypes.c_int(0)
    check(load().rrlbm_device_count(ctypes.byref(n)))
    return n.value


def diag_flow_noalloc(rho, ux, uy, uz, n, stream=None):
    """rrlbm_diag_flow_noalloc: the diag_flow sums of caller-owned FP64 device
    fields (device addresses as integers) of n cells, with no device memory
    allocated by the library; returns the rrlbm_diag as a dict."""
    d = Diag()
    lib = load()
    D = ctypes.POINTER(ctypes.c_double)
    check(lib.rrlbm_diag_flow_noalloc(ctypes.cast(rho, D), ctypes.cast(ux, D), ctypes.cast(uy, D),
                                      ctypes.cast(uz, D), int(n), ctypes.byref(d), stream))
    return d.as_dict()


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Solver:
    """One simulation context on the current GPU."""

    def __init__(self, params=None, **overrides):
        self._lib = load()
        p = params if params is not None else default_params()
        if overrides:
            q = Params.from_buffer_copy(p)
            for k, v in overrides.items():
                if k in ("force", "lid_u"):
                    arr = getattr(q, k)
                    for i in range(3):
                        arr[i] = float(v[i])
                else:
                    setattr(q, k, v)
            p = q
        h = ctypes.c_void_p()
        check(self._lib.rrlbm_create(ctypes.byref(p), ctypes.byref(h)))
        self._h = h
        self.params = Params()
        self._call("rrlbm_get_params", ctypes.byref(self.params))
        self.lx, self.ly, self.lz = self.params.lx, self.params.ly, self.params.lz
        self.npop = self.params.npop
        self.n = self.lx * self.ly * self.lz

    # -- plumbing ---------------------------------------------------------
    def _call(self, name, *args):
        if self._h is None:
            raise RRLBMError(-1, "solver is closed")
        check(getattr(self._lib, name)(self._h, *args), self._h)

    def close(self):
        if getattr(self, "_h", None) is not None:
            self._lib.rrlbm_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self):
        self._call("rrlbm_synchronize")

    def memory_bytes(self):
        b = ctypes.c_size_t(0)
        self._call("rrlbm_memory_bytes", ctypes.byref(b))
        return b.value

    # -- fields -----------------------------------------------------------
    def _size(self, field):
        return self.npop * self.n if field in (Field.F, Field.F_TEMP) else self.n

    def upload(self, field, data):
        a = np.ascontiguousarray(data, dtype=np.float64).ravel()
        self._call("rrlbm_upload", int(field), _dptr(a), a.size)

    def download(self, field):
        a = np.empty(self._size(field), dtype=np.float64)
        self._call("rrlbm_download", int(field), _dptr(a), a.size)
        if field in (Field.F, Field.F_TEMP):
            return a.reshape(self.npop, self.n)
        return a

    def macroscopic(self):
        """(rho, ux, uy, uz) as flat FP64 arrays."""
        return tuple(self.download(f) for f in (Field.RHO, Field.UX, Field.UY, Field.UZ))

    # -- reference routines (RR27V33_code_DNS_LES) ------------------------
    def init_arrays(self, ustar, ystar):
        """init_arrays, initial.cu:268-325."""
        self._call("rrlbm_init_arrays", float(ustar), float(ystar))

    def forcing(self):
        """forcing, macrovar.cu:63-74."""
        self._call("rrlbm_forcing")

    def forcing_GP(self, istep, rand_phase):
        """forcing_GP, macrovar.cu:76-114."""
        self._call("rrlbm_forcing_gp", int(istep), float(rand_phase))

    def LES_WALE(self):
        """LES_WALE, LES.cu:4-98."""
        self._call("rrlbm_les_wale")

    def collision_BGK(self):
        """collision_BGK, BGK/collision.cu:3-33 (collision_model = 1)."""
        self._call("rrlbm_collision_bgk")

    def collision_RR(self):
        """collision_RR, collision.cu:4-119."""
        self._call("rrlbm_collision_rr")

    def streaming(self):
        """streaming, streaming.cu:3-24."""
        self._call("rrlbm_streaming")

    def BB(self):
        """BB, streaming.cu:26-55."""
        self._call("rrlbm_bb")

    def copy_f(self):
        """The f_temp -> f copy of main.cu:64 (a buffer swap)."""
        self._call("rrlbm_swap_f")

    def macrovar(self):
        """macrovar, macrovar.cu:4-31."""
        self._call("rrlbm_macrovar")

    def RRBC(self):
        """RRBC, BC.cu:4-126."""
        self._call("rrlbm_rrbc")

    def FDBC(self):
        """FDBC, BC.cu:129-309 (wall velocities extrapolated before the differences)."""
        self._call("rrlbm_fdbc")

    def macrovar_BN(self):
        """macrovar_BN, macrovar.cu:33-62."""
        self._call("rrlbm_macrovar_bn")

    def step_reference(self, nsteps=1):
        """nsteps of main.cu:50-84 with the drop-in kernels."""
        self._call("rrlbm_step_reference", int(nsteps))

    # -- initial conditions -------------------------------------------------
    def init_rest(self):
        self._call("rrlbm_init_rest")

    def init_taylor_green_2d(self, u0):
        self._call("rrlbm_init_taylor_green_2d", float(u0))

    def init_taylor_green_3d(self, u0):
        self._call("rrlbm_init_taylor_green_3d", float(u0))

    def init_channel(self, u_tau, visc, seed, amp_streak=0.1, amp_noise=0.01):
        self._call("rrlbm_init_channel", float(u_tau), float(visc), ctypes.c_uint64(seed),
                   float(amp_streak), float(amp_noise))

    def init_equilibrium(self):
        self._call("rrlbm_init_equilibrium")

    # -- fused path ---------------------------------------------------------
    def fused_pack(self):
        self._call("rrlbm_fused_pack")

    def fused_step(self, nsteps=1):
        self._call("rrlbm_fused_step", int(nsteps))

    def fused_step_gp(self, istep, rand_phase):
        """One fused step whose collision is step istep under forcing_GP
        (macrovar.cu:76-114), the force computed in the kernel."""
        self._call("rrlbm_fused_step_gp", int(istep), float(rand_phase))

    def fused_step_timed(self, nsteps):
        ms = ctypes.c_double(0.0)
        self._call("rrlbm_fused_step_timed", int(nsteps), ctypes.byref(ms))
        return ms.value

    def fused_unpack(self):
        self._call("rrlbm_fused_unpack")

    def fused_diag(self):
        d = Diag()
        self._call("rrlbm_fused_diag", ctypes.byref(d))
        return d.as_dict()

    # -- diagnostics and files ----------------------------------------------
    def diag_flow(self):
        """diag_flow sums, output.cu:3-66."""
        d = Diag()
        self._call("rrlbm_diag_flow", ctypes.byref(d))
        return d.as_dict()

    def output_profile(self):
        """output_profile statistics, output.cu:142-225: array [LY, 10]."""
        a = np.empty(10 * self.ly, dtype=np.float64)
        self._call("rrlbm_plane_profiles", _dptr(a), a.size)
        return a.reshape(self.ly, 10)

    def write_flow(self, directory, istep):
        self._call("rrlbm_write_flow", directory.encode(), int(istep))

    def write_f(self, directory, istep):
        self._call("rrlbm_write_f", directory.encode(), int(istep))

    def read_restart(self, directory, istep):
        self._call("rrlbm_read_restart", directory.encode(), int(istep))

    def release_reference(self):
        self._call("rrlbm_release_reference")

    def graph_capture(self, nsteps):
        """Capture nsteps fused steps into a HIP graph (after fused_pack)."""
        self._call("rrlbm_graph_capture", int(nsteps))

    def graph_launch(self, times=1):
        self._call("rrlbm_graph_launch", int(times))

    def graph_launch_timed(self, times=1):
        """Replay the graph `times` times; returns the elapsed ms (HIP events)."""
        ms = ctypes.c_double(0.0)
        self._call("rrlbm_graph_launch_timed", int(times), ctypes.byref(ms))
        return ms.value

    def autotune(self, trial_steps=20):
        """Pick the fastest population layout and launch shape for this
        configuration (before fused_pack); returns the choice as a dict."""
        t = Tuning()
        self._call("rrlbm_autotune", int(trial_steps), ctypes.byref(t))
        self._call("rrlbm_get_params", ctypes.byref(self.params))
        return t.as_dict()


class _Borrowed(Solver):
    """A Solver view of one batch instance (owned by the batch)."""

    def __init__(self, lib, handle):
        self._lib = lib
        self._h = handle
        self.params = Params()
        self._call("rrlbm_get_params", ctypes.byref(self.params))
        self.lx, self.ly, self.lz = self.params.lx, self.params.ly, self.params.lz
        self.npop = self.params.npop
        self.n = self.lx * self.ly * self.lz

    def close(self):
        self._h = None


class Batch:
    """Ensemble of periodic D3Q27 boxes stepped by one launch per step."""

    def __init__(self, instances, storage=Storage.F32, keep_reference=True):
        self._lib = load()
        arr = (Instance * len(instances))()
        for k, (lx, ly, lz, tau, u0) in enumerate(instances):
            arr[k].lx, arr[k].ly, arr[k].lz, arr[k].tau, arr[k].u0 = lx, ly, lz, tau, u0
        h = ctypes.c_void_p()
        check_batch(self._lib.rrlbm_batch_create(arr, len(instances), int(storage), int(keep_reference),
                                                 ctypes.byref(h)))
        self._h = h
        self.count = len(instances)

    def _call(self, name, *args):
        if self._h is None:
            raise RRLBMError(-1, "batch is closed")
        check_batch(getattr(self._lib, name)(self._h, *args), self._h)

    def close(self):
        if getattr(self, "_h", None) is not None:
            self._lib.rrlbm_batch_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def cells(self):
        n = ctypes.c_int64(0)
        self._call("rrlbm_batch_cells", ctypes.byref(n))
        return n.value

    def init_taylor_green_3d(self):
        self._call("rrlbm_batch_init_taylor_green_3d")

    def set_grid(self, wg_per_cu):
        """0: one workgroup per tile; k: persistent grid of k workgroups per CU."""
        self._call("rrlbm_batch_set_grid", int(wg_per_cu))

    def autotune(self, trial_steps=10):
        """Time the batched launch's grids (before init); returns (wg_per_cu, ms per step)."""
        k = ctypes.c_int(0)
        ms = ctypes.c_double(0.0)
        self._call("rrlbm_batch_autotune", int(trial_steps), ctypes.byref(k), ctypes.byref(ms))
        return k.value, ms.value

    def step(self, nsteps):
        self._call("rrlbm_batch_step", int(nsteps))

    def step_sequential(self, nsteps):
        self._call("rrlbm_batch_step_sequential", int(nsteps))

    def graph_capture(self, nsteps):
        self._call("rrlbm_batch_graph_capture", int(nsteps))

    def graph_launch(self, times=1):
        self._call("rrlbm_batch_graph_launch", int(times))

    def time(self, mode, n):
        """mode 0: n batched steps, 1: n sequential steps, 2: n graph replays; returns ms."""
        ms = ctypes.c_double(0.0)
        self._call("rrlbm_batch_time", int(mode), int(n), ctypes.byref(ms))
        return ms.value

    def diag(self, k):
        d = Diag()
        self._call("rrlbm_batch_diag", int(k), ctypes.byref(d))
        return d.as_dict()

    def instance(self, k):
        h = ctypes.c_void_p()
        self._call("rrlbm_batch_context", int(k), ctypes.byref(h))
        return _Borrowed(self._lib, h)
