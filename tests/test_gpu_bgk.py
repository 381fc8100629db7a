"""The BGK baseline operator (the reference's BGK solver, BGK/) against the
oracle's restatement of BGK/collision.cu and BGK/macrovar.cu.

The oracle keeps the BGK solver's own variables (shifted populations f - w,
density deviation rho - 1); the library keeps its reference-layout
convention (f unshifted, rho = 1 + deviation), so the comparison converts.
FP64: drop-in kernels in the BGK/main.cu order to 1e-13, the fused step to
1e-12; FP32 and FP16 storage against the FP64 fused run."""
import numpy as np
import pytest

import oracle
from helpers import maxabs

pytestmark = pytest.mark.gpu


def bgk_state(lx, ly, lz, seed, force):
    rng = np.random.default_rng(seed)
    g = oracle.Fields(oracle.D3Q27_CUDA, lx, ly, lz)
    c, w, _ = oracle.lattice(oracle.D3Q27_CUDA)
    n = g.n
    rho = 0.01 * rng.standard_normal(n)
    u = 0.04 * rng.uniform(-1, 1, (3, n))
    eu = 3.0 * (c.astype(float) @ u)
    uv = 3.0 * (u * u).sum(0)
    g.f[:] = w[:, None] * (rho[None, :] + eu + 0.5 * (eu * eu - uv[None, :]))
    g.f += 2e-3 * w[:, None] * rng.standard_normal((27, n))
    g.set_force(force)
    g.macrovar_bgk()
    return g, w


def upload_bgk(s, g, w, rr):
    s.upload(rr.Field.F, g.f + w[:, None])
    s.upload(rr.Field.RHO, g.rho + 1.0)
    for fld, a in ((rr.Field.UX, g.ux), (rr.Field.UY, g.uy), (rr.Field.UZ, g.uz), (rr.Field.FORCEX, g.fx),
                   (rr.Field.FORCEY, g.fy), (rr.Field.FORCEZ, g.fz)):
        s.upload(fld, a)


FORCE = (3e-5, -1e-5, 2e-6)


@pytest.mark.parametrize("lx", [20, 64])
def test_bgk_reference_order_matches_oracle(rr, lx):
    g, w = bgk_state(lx, 12, 6, 4, FORCE)
    tau, steps = 0.57, 8
    with rr.Solver(lx=lx, ly=12, lz=6, storage=0, tau=tau, force=FORCE, wall_y=1, wall_rr=0,
                   collision_model=1) as s:
        upload_bgk(s, g, w, rr)
        s.step_reference(steps)
        f = s.download(rr.Field.F)
        rho, ux, uy, uz = s.macroscopic()
    g.step_bgk(steps, tau, walls=(0, 1, 0))
    assert maxabs(f - w[:, None], g.f) < 1e-13
    assert maxabs(rho - 1.0, g.rho) < 1e-13
    for a, b in ((ux, g.ux), (uy, g.uy), (uz, g.uz)):
        assert maxabs(a, b) < 1e-13


@pytest.mark.parametrize("lx", [20, 64])
def test_bgk_fused_matches_oracle(rr, lx):
    g, w = bgk_state(lx, 12, 6, 5, FORCE)
    tau, steps = 0.61, 10
    res = {}
    for st in (0, 1, 2):
        with rr.Solver(lx=lx, ly=12, lz=6, storage=st, tau=tau, force=FORCE, wall_y=1, wall_rr=0,
                       collision_model=1) as s:
            upload_bgk(s, g, w, rr)
            s.fused_pack()
            s.fused_step(steps - 1)
            s.fused_unpack()
            res[st] = (s.download(rr.Field.F), s.macroscopic(), s.fused_diag())
    ref = g.copy()
    ref.step_bgk(steps, tau, walls=(0, 1, 0))
    f64, m64, d64 = res[0]
    assert maxabs(f64 - w[:, None], ref.f) < 1e-12
    assert maxabs(m64[1], ref.ux) < 1e-12 and maxabs(m64[2], ref.uy) < 1e-12
    assert abs(d64["mass_excess"] - ref.rho.sum()) < 1e-10
    umax = np.abs(m64[1]).max()
    assert maxabs(res[1][1][1], m64[1]) < 1e-5 * umax
    assert maxabs(res[2][1][1], m64[1]) < 3e-2 * umax


def test_bgk_parameter_checks(rr):
    for bad in (dict(use_les=1), dict(wall_rr=1), dict(npop=19), dict(collision_model=2)):
        kw = dict(lx=8, ly=8, lz=8, wall_y=1, wall_rr=0, collision_model=1)
        kw.update(bad)
        with pytest.raises(rr.RRLBMError) as ei:
            rr.Solver(**kw)
        assert ei.value.status in (1, 6)
    with rr.Solver(lx=8, ly=8, lz=8) as s:
        with pytest.raises(rr.RRLBMError) as ei:
            s.collision_BGK()
        assert ei.value.status == 3
