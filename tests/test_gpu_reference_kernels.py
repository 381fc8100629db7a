"""GPU parity of every drop-in kernel (FP64, reference layout) against the oracle.

Each test puts the same random state on the GPU and in the oracle, runs one
reference routine on both and compares the fields.  Tolerance: 1e-13 absolute
on populations (O(1) values) and velocities -- FP64 rounding differences
between the moment-space GPU formulation and the per-direction oracle only.
"""
import numpy as np
import pytest

import oracle
from helpers import channel_state, maxabs, random_state, upload_state

pytestmark = pytest.mark.gpu

TOL = 1e-13
GRIDS = [(12, 10, 8), (7, 5, 6)]


def solver_for(rr, kind, dims, **kw):
    q = 19 if kind == oracle.D3Q19 else 27
    return rr.Solver(lx=dims[0], ly=dims[1], lz=dims[2], npop=q, storage=0, **kw)


@pytest.mark.parametrize("kind", [oracle.D3Q27_CUDA, oracle.D3Q19])
@pytest.mark.parametrize("dims", GRIDS)
@pytest.mark.parametrize("use_les", [0, 1])
def test_collision_rr(rr, kind, dims, use_les):
    if use_les and kind == oracle.D3Q19:
        pytest.skip("WALE is D3Q27 only")
    g = random_state(kind, *dims, seed=1)
    tau = 0.56
    with solver_for(rr, kind, dims, tau=tau, use_les=use_les) as s:
        upload_state(s, g)
        s.collision_RR()
        got = s.download(rr.Field.F)
    g.collision(tau, use_les=bool(use_les))
    assert maxabs(got, g.f) < TOL


@pytest.mark.parametrize("kind", [oracle.D3Q27_CUDA, oracle.D3Q19])
@pytest.mark.parametrize("dims", GRIDS)
def test_streaming_bb_macrovar(rr, kind, dims):
    g = random_state(kind, *dims, seed=2)
    with solver_for(rr, kind, dims) as s:
        upload_state(s, g)
        s.streaming()
        s.BB()
        s.copy_f()
        s.macrovar()
        got_f = s.download(rr.Field.F)
        got_post = s.download(rr.Field.F_TEMP)
        got_m = s.macroscopic()
    g.streaming()
    g.bb(walls=(0, 1, 0))
    g.swap()
    g.macrovar()
    assert maxabs(got_f, g.f) == 0.0          # pure data movement: bit-exact
    assert maxabs(got_post, g.ftmp) == 0.0
    for a, b in zip(got_m, (g.rho, g.ux, g.uy, g.uz)):
        assert maxabs(a, b) < TOL


@pytest.mark.parametrize("kind", [oracle.D3Q27_CUDA, oracle.D3Q19])
@pytest.mark.parametrize("walls,lid_face", [((0, 1, 0), -1), ((1, 1, 0), 3), ((1, 1, 1), 3)])
def test_rrbc_macrovar_bn(rr, kind, walls, lid_face):
    dims = (9, 8, 6)
    lid_u = (0.05, 0.0, 0.02)
    g = random_state(kind, *dims, seed=3)
    kw = dict(wall_x=walls[0], wall_y=walls[1], wall_z=walls[2], lid_face=lid_face, lid_u=lid_u)
    with solver_for(rr, kind, dims, **kw) as s:
        upload_state(s, g)
        s.streaming()
        s.BB()
        s.copy_f()
        s.macrovar()
        s.RRBC()
        s.macrovar_BN()
        got_f = s.download(rr.Field.F)
        got_m = s.macroscopic()
    g.streaming()
    g.bb(walls=walls, lid_face=lid_face, lid_u=lid_u)
    g.swap()
    g.macrovar()
    g.rrbc(walls=walls, lid_face=lid_face, lid_u=lid_u)
    g.macrovar(wall_only=1, walls=walls)
    assert maxabs(got_f, g.f) < TOL
    for a, b in zip(got_m, (g.rho, g.ux, g.uy, g.uz)):
        assert maxabs(a, b) < TOL


@pytest.mark.parametrize("dims", GRIDS)
def test_les_wale(rr, dims):
    g = random_state(oracle.D3Q27_CUDA, *dims, seed=4)
    with solver_for(rr, oracle.D3Q27_CUDA, dims, use_les=1) as s:
        upload_state(s, g)
        s.LES_WALE()
        got = s.download(rr.Field.TAUT)
    g.les_wale(cw=0.5, guard=1)
    assert np.allclose(got, g.taut, rtol=1e-12, atol=1e-15)


def test_forcing_and_gp(rr):
    dims = (24, 16, 8)
    g = oracle.Fields(oracle.D3Q27_CUDA, *dims)
    fin = 3.1640625e-07
    with solver_for(rr, oracle.D3Q27_CUDA, dims, force=(fin, 0, 0)) as s:
        s.forcing()
        got = [s.download(f) for f in (rr.Field.FORCEX, rr.Field.FORCEY, rr.Field.FORCEZ)]
        g.forcing(fin)
        for a, b in zip(got, (g.fx, g.fy, g.fz)):
            assert maxabs(a, b) == 0.0
        for istep, phase in ((1, 0.0), (1999, 0.37), (2500, 0.91)):
            s.forcing_GP(istep, phase)
            got = [s.download(f) for f in (rr.Field.FORCEX, rr.Field.FORCEY, rr.Field.FORCEZ)]
            g.forcing_gp(fin, istep, phase)
            for a, b in zip(got, (g.fx, g.fy, g.fz)):
                assert np.allclose(a, b, rtol=1e-13, atol=1e-22)


def test_init_arrays(rr):
    dims = (16, 12, 8)
    visc = 0.000622
    u0 = 2.0 * 180.0 * visc / dims[1]
    ystar = visc / u0
    fin = 2.0 * u0 * u0 / dims[1]
    g = oracle.Fields(oracle.D3Q27_CUDA, *dims)
    with solver_for(rr, oracle.D3Q27_CUDA, dims, force=(fin, 0, 0)) as s:
        s.init_arrays(u0, ystar)
        got_f = s.download(rr.Field.F)
        got_m = s.macroscopic()
    g.init_arrays(u0, ystar, fin)
    assert maxabs(got_f, g.f) < 1e-16
    for a, b in zip(got_m, (g.rho, g.ux, g.uy, g.uz)):
        assert np.allclose(a, b, rtol=1e-14, atol=1e-17)


def test_step_reference_sequence(rr):
    """Ten steps of the main.cu sequence through rrlbm_step_reference vs the oracle."""
    dims = (12, 10, 8)
    g = random_state(oracle.D3Q27_CUDA, *dims, seed=5, force=(2e-5, 0.0, 0.0))
    tau = 0.52
    with solver_for(rr, oracle.D3Q27_CUDA, dims, tau=tau, use_les=1) as s:
        upload_state(s, g)
        s.step_reference(10)
        got_f = s.download(rr.Field.F)
        got_m = s.macroscopic()
    g.step(10, tau, use_les=True)
    assert maxabs(got_f, g.f) < 1e-12
    for a, b in zip(got_m, (g.rho, g.ux, g.uy, g.uz)):
        assert maxabs(a, b) < 1e-12


@pytest.mark.parametrize("les", [False, True])
@pytest.mark.parametrize("lx", [12, 64])
def test_fdbc_matches_oracle(rr, les, lx):
    """FDBC (BC.cu:129-309) alone and inside the reference-order step
    (wall_rr = 2), against the oracle's restatement."""
    g = channel_state(oracle.D3Q27_CUDA, lx, 10, 6, u_tau=0.01, visc=0.002, seed=9, force=(2e-5, 1e-6, 0))
    g.taut[:] = 0.02 * np.random.default_rng(1).uniform(0, 1, g.n)
    tau = 0.5 + 3 * 0.002
    with rr.Solver(lx=lx, ly=10, lz=6, storage=0, tau=tau, force=(2e-5, 1e-6, 0), wall_y=1, wall_rr=2,
                   use_les=int(les)) as s:
        upload_state(s, g)
        s.FDBC()
        f1 = s.download(rr.Field.F)
        u1 = s.macroscopic()[1:]
        s.step_reference(5)
        f2 = s.download(rr.Field.F)
        m2 = s.macroscopic()
    g.fdbc(tau, use_les=les)
    assert maxabs(f1, g.f) < 1e-14
    for a, b in zip(u1, (g.ux, g.uy, g.uz)):
        assert maxabs(a, b) < 1e-16
    g.step(5, tau, use_les=les, walls=(0, 1, 0), wall_rr=2)
    assert maxabs(f2, g.f) < 1e-13
    for a, b in zip(m2, (g.rho, g.ux, g.uy, g.uz)):
        assert maxabs(a, b) < 1e-13


def test_fdbc_parameter_checks(rr):
    for bad in (dict(npop=19), dict(wall_x=1), dict(ly=3), dict(lid_face=2, lid_u=(0.01, 0, 0)), dict(rr_order=6)):
        kw = dict(lx=8, ly=8, lz=4, wall_y=1, wall_rr=2)
        kw.update(bad)
        with pytest.raises(rr.RRLBMError) as ei:
            rr.Solver(**kw)
        assert ei.value.status == 6, bad
    with rr.Solver(lx=8, ly=8, lz=4, wall_y=1, wall_rr=2, use_les=1) as s:
        with pytest.raises(rr.RRLBMError) as ei:
            s.fused_pack()
        assert ei.value.status == 6
    with rr.Solver(lx=8, ly=8, lz=4, wall_y=1, wall_rr=1) as s:
        with pytest.raises(rr.RRLBMError):
            rr.Solver(lx=8, ly=8, lz=4, wall_y=1, wall_rr=3)
