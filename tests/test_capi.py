"""The C ABI library loads and exports every entry point include/rrlbm.h
declares; host-only entry points (no GPU work) behave as documented."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle
from conftest import REPO

HEADER = os.path.join(REPO, "include", "rrlbm.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+char\s*\*|int)\s+(rrlbm_\w+)\s*\(", text, re.M)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert len(names) >= 39
    for must in ("rrlbm_collision_rr", "rrlbm_streaming", "rrlbm_bb", "rrlbm_macrovar", "rrlbm_rrbc",
                 "rrlbm_macrovar_bn", "rrlbm_les_wale", "rrlbm_forcing", "rrlbm_forcing_gp", "rrlbm_init_arrays",
                 "rrlbm_fused_step", "rrlbm_diag_flow", "rrlbm_plane_profiles"):
        assert must in names


def test_library_exports_every_declared_symbol(rr):
    lib = ctypes.CDLL(rr.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    # the Python mirror binds all of them too
    from rrlbm._lib import SIGNATURES
    bound = set(SIGNATURES) | {"rrlbm_version", "rrlbm_last_error", "rrlbm_batch_last_error"}
    assert set(declared_functions()) <= bound


def test_params_default_match_reference_globals(rr):
    p = rr.default_params()
    visc, Re, ly = 0.000622, 180.0, 64
    u0 = 2.0 * Re * visc / ly                      # globals.cu:59
    assert (p.lx, p.ly, p.lz, p.npop, p.use_les) == (64, 64, 64, 27, 0)   # parameters.h:6-13
    assert p.tau == 3.0 * visc + 0.5               # globals.cu:60
    assert p.force[0] == 2.0 * u0 * u0 / ly        # globals.cu:63
    assert p.cw == 0.5                             # LES.cu:12
    assert p.wall_y == 1 and p.wall_rr == 1 and p.lid_face == -1


@pytest.mark.parametrize("q,kind", [(27, oracle.D3Q27_CUDA), (19, oracle.D3Q19)])
def test_lattice_ordering_matches_oracle(rr, q, kind):
    c, w, opp = rr.lattice(q)
    C, W, O = oracle.lattice(kind)
    assert (c == C).all() and np.allclose(w, W, rtol=1e-15) and (opp == O).all()


def test_invalid_parameters_are_reported_not_fatal(rr):
    """Argument validation runs before any device work, with a message."""
    lib = rr.load()
    h = ctypes.c_void_p()
    for bad in (dict(npop=15), dict(tau=0.4), dict(use_les=1, npop=19), dict(storage=7), dict(lx=0),
                dict(lid_face=3, wall_y=0), dict(f16_rounding=2)):
        p = rr.default_params(**bad)
        st = lib.rrlbm_create(ctypes.byref(p), ctypes.byref(h))
        assert st in (1, 6)
        assert lib.rrlbm_last_error(None)
    # limits and option combinations, with the exact status: 1 invalid, 6 unsupported
    cases = [(dict(lx=2048, ly=1024, lz=1024), 1),                      # 2^31 cells
             (dict(lx=1024, ly=1024, lz=1000, storage=0), 1),           # a population array over 4 GiB
             (dict(xcd_remap=-2), 1), (dict(xcd_remap=65537), 1), (dict(cells_per_thread=3), 1), (dict(persistent_wg_per_cu=65), 1),
             (dict(collision_model=2), 1), (dict(collision_model=1, wall_rr=1), 6), (dict(collision_model=1, npop=19), 6),
             (dict(rr_order=5), 1), (dict(rr_order=6, npop=19), 6), (dict(rr_order=6, use_les=1), 6),
             (dict(wall_rr=3), 1), (dict(wall_rr=2, wall_y=1, wall_x=1), 6), (dict(wall_rr=2, wall_y=1, ly=3), 6),
             (dict(wall_rr=2, wall_y=1, npop=19), 6), (dict(wall_rr=2, wall_y=1, rr_order=6), 6),
             (dict(wall_y=1, ly=1), 1), (dict(lid_face=6), 1),
             # two steps per launch: the D3Q27 RR (orders 3, 6) and BGK steps without LES or FDBC,
             # FP32/FP64, lx a multiple of 64, ly of 4
             (dict(steps_per_pass=3), 1), (dict(steps_per_pass=-1), 1), (dict(pass_planes=-1), 1),
             (dict(steps_per_pass=2, storage=2), 6), (dict(steps_per_pass=2, lx=96), 6),
             (dict(steps_per_pass=2, ly=66), 6), (dict(steps_per_pass=2, use_les=1), 6),
             (dict(steps_per_pass=2, npop=19), 6), (dict(steps_per_pass=2, wall_rr=2), 6)]
    for bad, want in cases:
        p = rr.default_params(**bad)
        st = lib.rrlbm_create(ctypes.byref(p), ctypes.byref(h))
        assert st == want, (bad, st, lib.rrlbm_last_error(None))
    assert lib.rrlbm_lattice(20, None, None, None) == 1
