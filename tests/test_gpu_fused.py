"""GPU parity of the fused step against the oracle's reference-order step.

pack (the first collision) + (K-1) fused launches + unpack (the last
streaming / BB / macrovar / RRBC / macrovar_BN) must equal K steps of the
main.cu sequence (RR27V33_code_DNS_LES/main.cu:50-84) run by the oracle.
FP64 storage: 1e-12 absolute on populations and velocities.  FP32 / FP16
storage are checked against the FP64 run of the same kernels (relative errors
stated per test).
"""
import numpy as np
import pytest

import oracle
from helpers import channel_state, maxabs, random_state, upload_state

pytestmark = pytest.mark.gpu


def run_fused(rr, g, steps, **params):
    with rr.Solver(lx=g.lx, ly=g.ly, lz=g.lz, npop=g.q, **params) as s:
        upload_state(s, g)
        s.fused_pack()
        s.fused_step(steps - 1)
        s.fused_unpack()
        diag = s.fused_diag()
        return s.download(rr.Field.F), s.macroscopic(), diag


CASES = {
    # name: (kind, dims, params, oracle walls, lid)
    "channel_q27": (oracle.D3Q27_CUDA, (16, 12, 8), dict(wall_y=1, wall_rr=1), (0, 1, 0), -1),
    "periodic_q27": (oracle.D3Q27_CUDA, (10, 9, 7), dict(wall_y=0, wall_rr=0), (0, 0, 0), -1),
    "cavity_q27": (oracle.D3Q27_CUDA, (9, 8, 7), dict(wall_x=1, wall_y=1, wall_z=1, wall_rr=1, lid_face=3,
                                                       lid_u=(0.05, 0, 0)), (1, 1, 1), 3),
    "cavity_q19": (oracle.D3Q19, (9, 8, 7), dict(wall_x=1, wall_y=1, wall_z=1, wall_rr=1, lid_face=3,
                                                  lid_u=(0.05, 0, 0)), (1, 1, 1), 3),
    "periodic_q19": (oracle.D3Q19, (10, 9, 7), dict(wall_y=0, wall_rr=0), (0, 0, 0), -1),
    "cavity2d_bb_only_q27": (oracle.D3Q27_CUDA, (10, 8, 2), dict(wall_x=1, wall_y=1, wall_rr=0, lid_face=1,
                                                                  lid_u=(0, 0.05, 0)), (1, 1, 0), 1),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_fused_fp64_matches_reference_order(rr, name):
    kind, dims, params, walls, lid = CASES[name]
    force = (0.0, 0.0, 0.0) if lid >= 0 else (3e-5, 1e-6, -2e-6)
    g = random_state(kind, *dims, seed=11, force=force)
    tau = 0.53
    steps = 12
    lid_u = params.get("lid_u", (0.0, 0.0, 0.0))
    got_f, got_m, diag = run_fused(rr, g, steps, storage=0, tau=tau, force=force, **params)
    g.step(steps, tau, walls=walls, wall_rr=params.get("wall_rr", 0), lid_face=lid, lid_u=lid_u)
    assert maxabs(got_f, g.f) < 1e-12
    for a, b in zip(got_m, (g.rho, g.ux, g.uy, g.uz)):
        assert maxabs(a, b) < 1e-12
    ref = g.diag()
    assert abs(diag["mass_excess"] - ref[4]) < 1e-10
    assert abs(diag["kinetic_energy"] - ref[8]) < 1e-12 * max(1.0, abs(ref[8]))
    assert np.allclose(diag["momentum"], ref[5:8], rtol=1e-10, atol=1e-13)


@pytest.mark.parametrize("lagged", [0, 1])
def test_fused_les(rr, lagged):
    """WALE in the fused path: reference order (two launches) is exact; the
    one-launch lagged-gradient variant differs by the one-step lag only."""
    g = channel_state(oracle.D3Q27_CUDA, 16, 12, 8, u_tau=0.01, visc=0.002, seed=7, force=(2e-5, 0, 0))
    tau = 0.5 + 3 * 0.002
    steps = 10
    got_f, got_m, _ = run_fused(rr, g, steps, storage=0, tau=tau, force=(2e-5, 0, 0), use_les=1,
                                les_lagged=lagged)
    g.step(steps, tau, use_les=True, walls=(0, 1, 0), wall_rr=1)
    err = maxabs(got_f, g.f)
    if lagged == 0:
        assert err < 1e-12
    else:
        assert err < 1e-5  # lag of one step in tau_t only


def test_fused_fp32_fp16_vs_fp64(rr):
    """FP32 and FP16-shifted storage (FP32 arithmetic) against the FP64 run of
    the same kernels on a channel state, 50 steps."""
    g = channel_state(oracle.D3Q27_CUDA, 32, 24, 16, u_tau=0.006, visc=0.0016, seed=3, force=(1e-5, 0, 0))
    tau = 0.5 + 3 * 0.0016
    res = {}
    for st in (0, 1, 2):
        _, m, diag = run_fused(rr, g, 50, storage=st, tau=tau, force=(1e-5, 0, 0), wall_y=1, wall_rr=1)
        res[st] = (m, diag)
    umax = np.abs(res[0][0][1]).max()
    e32 = maxabs(res[1][0][1], res[0][0][1]) / umax
    e16 = maxabs(res[2][0][1], res[0][0][1]) / umax
    assert e32 < 2e-5, e32
    assert e16 < 2e-2, e16
    ke = res[0][1]["kinetic_energy"]
    assert abs(res[1][1]["kinetic_energy"] - ke) / ke < 1e-5
    assert abs(res[2][1]["kinetic_energy"] - ke) / ke < 2e-3


def test_pack_state_error_before_pack(rr):
    with rr.Solver(lx=8, ly=8, lz=8) as s:
        with pytest.raises(rr.RRLBMError) as ei:
            s.fused_step(1)
        assert ei.value.status == 3


@pytest.mark.parametrize("walls", [(0, 1, 0), (1, 1, 0), (0, 0, 0)])
def test_two_cells_per_thread_matches_reference_order(rr, walls):
    """The pair kernel (two x-adjacent cells per thread) gives the reference
    order's result in FP64 and agrees with the one-cell kernel in FP32/FP16."""
    g = random_state(oracle.D3Q27_CUDA, 12, 10, 6, seed=31, force=(3e-5, 0.0, 0.0))
    tau = 0.54
    kw = dict(tau=tau, force=(3e-5, 0, 0), wall_x=walls[0], wall_y=walls[1], wall_z=walls[2],
              wall_rr=1 if any(walls) else 0)
    got_f, got_m, _ = run_fused(rr, g, 9, storage=0, cells_per_thread=2, **kw)
    ref = g.copy()
    ref.step(9, tau, walls=walls, wall_rr=kw["wall_rr"])
    assert maxabs(got_f, ref.f) < 1e-12
    for st, tol in ((1, 1e-6), (2, 1e-3)):
        f1, _, _ = run_fused(rr, g, 9, storage=st, cells_per_thread=1, **kw)
        f2, _, _ = run_fused(rr, g, 9, storage=st, cells_per_thread=2, **kw)
        assert maxabs(f1, f2) < tol
