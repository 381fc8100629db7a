"""The RR operator with the fourth- to sixth-order Hermite terms
(rrlbm_params.rr_order = 6; the terms the reference keeps commented out in
RR27V33_code_DNS_LES/collision.cu:48-49,62-67,95-104,112-116 and
BC.cu:37-38,50-54,79-88,97-101) against the oracle's restatement at order 6.

FP64: drop-in kernels in the main.cu order and the fused step to 1e-12 (the
same bars as the third-order operator); FP32 / FP16 storage against the FP64
run of the same kernels."""
import numpy as np
import pytest

import oracle
from helpers import channel_state, maxabs, random_state, upload_state

pytestmark = pytest.mark.gpu

CASES = {
    "channel": ((16, 12, 8), dict(wall_y=1, wall_rr=1), (0, 1, 0), -1),
    "channel_rows64": ((64, 6, 5), dict(wall_y=1, wall_rr=1), (0, 1, 0), -1),
    # 73 728 cells = 288 blocks: more than a persistent grid of one workgroup
    # per compute unit has on the 256 CUs of an MI355X, so that grid strides
    "channel_rows128_288blocks": ((128, 24, 24), dict(wall_y=1, wall_rr=1), (0, 1, 0), -1),
    "periodic": ((10, 9, 7), dict(wall_y=0, wall_rr=0), (0, 0, 0), -1),
    "cavity": ((9, 8, 7), dict(wall_x=1, wall_y=1, wall_z=1, wall_rr=1, lid_face=3, lid_u=(0.05, 0, 0)),
               (1, 1, 1), 3),
}


def case_state(name, seed):
    dims, params, walls, lid = CASES[name]
    force = (0.0, 0.0, 0.0) if lid >= 0 else (3e-5, 1e-6, -2e-6)
    g = random_state(oracle.D3Q27_CUDA, *dims, seed=seed, amp=3e-3, umax=0.1, force=force)
    return g, dims, params, walls, lid, force


@pytest.mark.parametrize("name", sorted(CASES))
def test_rr6_reference_order_matches_oracle(rr, name):
    g, dims, params, walls, lid, force = case_state(name, 21)
    tau, steps = 0.56, 6
    with rr.Solver(lx=dims[0], ly=dims[1], lz=dims[2], storage=0, tau=tau, force=force, rr_order=6,
                   **params) as s:
        upload_state(s, g)
        s.step_reference(steps)
        f = s.download(rr.Field.F)
        m = s.macroscopic()
    with oracle.rr_order(6):
        g.step(steps, tau, walls=walls, wall_rr=params["wall_rr"], lid_face=lid,
               lid_u=params.get("lid_u", (0.0, 0.0, 0.0)))
    assert maxabs(f, g.f) < 1e-13
    for a, b in zip(m, (g.rho, g.ux, g.uy, g.uz)):
        assert maxabs(a, b) < 1e-13


@pytest.mark.parametrize("name", sorted(CASES))
def test_rr6_fused_matches_oracle(rr, name):
    g, dims, params, walls, lid, force = case_state(name, 22)
    tau, steps = 0.53, 10
    with rr.Solver(lx=dims[0], ly=dims[1], lz=dims[2], storage=0, tau=tau, force=force, rr_order=6,
                   **params) as s:
        upload_state(s, g)
        s.fused_pack()
        s.fused_step(steps - 1)
        s.fused_unpack()
        f = s.download(rr.Field.F)
        m = s.macroscopic()
    g3 = g.copy()
    with oracle.rr_order(6):
        g.step(steps, tau, walls=walls, wall_rr=params["wall_rr"], lid_face=lid,
               lid_u=params.get("lid_u", (0.0, 0.0, 0.0)))
    assert maxabs(f, g.f) < 1e-12
    for a, b in zip(m, (g.rho, g.ux, g.uy, g.uz)):
        assert maxabs(a, b) < 1e-12
    # the third-order operator gives a different state from the same start
    g3.step(steps, tau, walls=walls, wall_rr=params["wall_rr"], lid_face=lid,
            lid_u=params.get("lid_u", (0.0, 0.0, 0.0)))
    assert maxabs(f, g3.f) > 1e-9


@pytest.mark.parametrize("name", ["channel", "channel_rows64", "channel_rows128_288blocks"])
def test_rr6_fused_force_x_matches_oracle(rr, name):
    """The step kernel for a force along x only (the channel's), which leaves
    out the y and z force terms of the sixth-order output, against the oracle:
    one workgroup per block and a persistent grid of one workgroup per compute
    unit (which strides over the blocks on the 288-block grid)."""
    dims, params, walls, lid = CASES[name]
    force = (3e-5, 0.0, 0.0)
    tau, steps = 0.53, 10
    for extra in ({}, {"persistent_wg_per_cu": 1}):
        g = random_state(oracle.D3Q27_CUDA, *dims, seed=23, amp=3e-3, umax=0.1, force=force)
        with rr.Solver(lx=dims[0], ly=dims[1], lz=dims[2], storage=0, tau=tau, force=force, rr_order=6,
                       **params, **extra) as s:
            upload_state(s, g)
            s.fused_pack()
            s.fused_step(steps - 1)
            s.fused_unpack()
            f = s.download(rr.Field.F)
            m = s.macroscopic()
        with oracle.rr_order(6):
            g.step(steps, tau, walls=walls, wall_rr=params["wall_rr"], lid_face=lid)
        assert maxabs(f, g.f) < 1e-12
        for a, b in zip(m, (g.rho, g.ux, g.uy, g.uz)):
            assert maxabs(a, b) < 1e-12


@pytest.mark.parametrize("lx", [32, 64])
def test_rr6_fp32_fp16_vs_fp64(rr, lx):
    g = channel_state(oracle.D3Q27_CUDA, lx, 24, 16, u_tau=0.006, visc=0.0016, seed=3, force=(1e-5, 0, 0))
    tau = 0.5 + 3 * 0.0016
    res = {}
    for st in (0, 1, 2):
        with rr.Solver(lx=lx, ly=24, lz=16, storage=st, tau=tau, force=(1e-5, 0, 0), wall_y=1, wall_rr=1,
                       rr_order=6) as s:
            upload_state(s, g)
            s.fused_pack()
            s.fused_step(49)
            s.fused_unpack()
            res[st] = s.macroscopic()
    umax = np.abs(res[0][1]).max()
    e32, e16 = maxabs(res[1][1], res[0][1]) / umax, maxabs(res[2][1], res[0][1]) / umax
    print("RR6 FP32/FP16 vs FP64 (lx %d): u %.2e / %.2e" % (lx, e32, e16))
    assert e32 < 5e-6      # measured 1.4e-6
    assert e16 < 3e-3      # measured 7.6e-4 .. 9.1e-4


def test_rr6_equilibrium_initialisation(rr):
    rng = np.random.default_rng(5)
    lx, ly, lz = 8, 6, 4
    g = oracle.Fields(oracle.D3Q27_CUDA, lx, ly, lz)
    g.rho[:] = 1 + 0.02 * rng.standard_normal(g.n)
    for a in (g.ux, g.uy, g.uz):
        a[:] = 0.1 * rng.uniform(-1, 1, g.n)
    with rr.Solver(lx=lx, ly=ly, lz=lz, storage=0, rr_order=6) as s:
        upload_state(s, g)
        s.init_equilibrium()
        f = s.download(rr.Field.F)
    with oracle.rr_order(6):
        g.init_equilibrium()
    assert maxabs(f, g.f) < 1e-15


def test_rr6_parameter_checks(rr):
    for bad, status in ((dict(npop=19), 6), (dict(use_les=1), 6), (dict(collision_model=1, wall_rr=0), 6),
                        (dict(rr_order=5), 1)):
        kw = dict(lx=8, ly=8, lz=8, wall_y=1, rr_order=6)
        kw.update(bad)
        with pytest.raises(rr.RRLBMError) as ei:
            rr.Solver(**kw)
        assert ei.value.status == status, (bad, ei.value)
    with rr.Solver(lx=8, ly=8, lz=8) as s:
        assert s.params.rr_order == 3
