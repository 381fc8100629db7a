"""Full-size configurations, checked through properties that do not depend on
the grid size (the oracle is too slow at these sizes): the fused path against
the drop-in kernels in the reference order at FP64, FP32 / FP16 against FP64
on the same grid, conservation and bit-identity of the launch forms.

Grids: the channel DNS 384 x 128 x 192 (9.4 M cells; also under the BGK
operator, the sixth-order RR terms and FDBC walls), the LES channel
192 x 64 x 96, the 128^3 cavity, the 64-instance Taylor-Green ensemble
(1.4 x 10^8 cells)."""
import numpy as np
import pytest

from helpers import maxabs

pytestmark = pytest.mark.gpu


def fused_vs_reference(rr, make, steps, **kw):
    s, _ = make(storage=0, **kw)
    with s:
        s.step_reference(steps)
        ref_f = s.download(rr.Field.F)
        ref_m = s.macroscopic()
    s, _ = make(storage=0, **kw)
    with s:
        s.fused_pack()
        s.fused_step(steps - 1)
        s.fused_unpack()
        f = s.download(rr.Field.F)
        m = s.macroscopic()
    return maxabs(f, ref_f), max(maxabs(a, b) for a, b in zip(m, ref_m))


def test_channel_dns_fused_equals_reference_order(rr):
    from rrlbm import configs
    ef, em = fused_vs_reference(rr, configs.channel, 4)
    assert ef < 1e-13 and em < 1e-13, (ef, em)


def test_channel_les_fused_equals_reference_order(rr):
    """The LES grid in the reference order (velocity pass + step: exact)."""
    from rrlbm import configs
    ef, em = fused_vs_reference(rr, configs.channel_les, 4, les_lagged=0)
    assert ef < 1e-13 and em < 1e-13, (ef, em)


@pytest.mark.parametrize("npop", [19, 27])
def test_cavity_fused_equals_reference_order(rr, npop):
    from rrlbm import configs
    ef, em = fused_vs_reference(rr, configs.cavity, 4, npop=npop)
    assert ef < 1e-13 and em < 1e-13, (ef, em)


@pytest.mark.parametrize("op", ["bgk", "rr6", "fdbc"])
def test_channel_operators_fused_equal_reference_order(rr, op):
    """The channel grid under the BGK operator (BGK/main.cu:49-70 order), the RR
    collision with the fourth- to sixth-order terms, and FDBC walls in place of
    RRBC (main.cu:76-77): the fused FP64 step against the drop-in kernels."""
    from rrlbm import configs
    make, kw = {"bgk": (configs.channel_bgk, {}), "rr6": (configs.channel, {"rr_order": 6}),
                "fdbc": (configs.channel, {"wall_rr": 2})}[op]
    ef, em = fused_vs_reference(rr, make, 4, **kw)
    print("%s channel 384x128x192, fused FP64 vs reference order: f %.2e, macroscopic %.2e" % (op, ef, em))
    assert ef < 1e-13 and em < 1e-13, (ef, em)


def test_channel_rr6_storage_precisions(rr):
    """The sixth-order RR step in FP32 / FP16 storage against FP64 after 100
    fused steps of the full channel."""
    from rrlbm import configs
    u = {}
    for st in (0, 1, 2):
        s, meta = configs.channel(storage=st, rr_order=6)
        with s:
            s.fused_pack()
            s.fused_step(99)
            d = s.fused_diag()
            s.fused_unpack()
            u[st] = (s.download(rr.Field.UX), d)
    umax = np.abs(u[0][0]).max()
    e32, e16 = maxabs(u[1][0], u[0][0]) / umax, maxabs(u[2][0], u[0][0]) / umax
    b32, b16 = (abs(u[st][1]["u_sum"][0] / u[0][1]["u_sum"][0] - 1) for st in (1, 2))
    print("channel RR6 FP32/FP16 vs FP64, 100 steps: u %.2e / %.2e, U_b %.2e / %.2e" % (e32, e16, b32, b16))
    # measured u 4.7e-6 / 1.2e-3, U_b 2.5e-7 / 4.7e-8 of the FP64 values
    assert e32 < 1.5e-5
    assert e16 < 4e-3
    assert b32 < 1e-6 and b16 < 4e-6


def test_channel_dns_storage_precisions(rr):
    """FP32 / FP16 against FP64 after 100 fused steps of the full channel."""
    from rrlbm import configs
    u = {}
    for st in (0, 1, 2):
        s, meta = configs.channel(storage=st)
        with s:
            s.fused_pack()
            s.fused_step(99)
            d = s.fused_diag()
            s.fused_unpack()
            u[st] = (s.download(rr.Field.UX), d)
    umax = np.abs(u[0][0]).max()
    e32, e16 = maxabs(u[1][0], u[0][0]) / umax, maxabs(u[2][0], u[0][0]) / umax
    b32, b16 = (abs(u[st][1]["u_sum"][0] / u[0][1]["u_sum"][0] - 1) for st in (1, 2))
    print("channel DNS FP32/FP16 vs FP64, 100 steps: u %.2e / %.2e, U_b %.2e / %.2e" % (e32, e16, b32, b16))
    # measured u 2.6e-6 / 1.3e-3, U_b 7.4e-7 / 1.1e-6 of the FP64 values
    assert e32 < 8e-6
    assert e16 < 2e-3
    assert b32 < 3e-6 and b16 < 4e-6


def test_channel_les_storage_precisions(rr):
    """The LES grid (192 x 64 x 96, WALE, lagged gradient) in FP32 / FP16
    storage against FP64 after 200 fused steps."""
    from rrlbm import configs
    u = {}
    for st in (0, 1, 2):
        s, meta = configs.channel_les(storage=st)
        with s:
            s.fused_pack()
            s.fused_step(199)
            d = s.fused_diag()
            s.fused_unpack()
            u[st] = (s.download(rr.Field.UX), d)
    umax = np.abs(u[0][0]).max()
    e32, e16 = maxabs(u[1][0], u[0][0]) / umax, maxabs(u[2][0], u[0][0]) / umax
    b32, b16 = (abs(u[st][1]["u_sum"][0] / u[0][1]["u_sum"][0] - 1) for st in (1, 2))
    print("channel LES FP32/FP16 vs FP64, 200 steps: u %.2e / %.2e, U_b %.2e / %.2e" % (e32, e16, b32, b16))
    # measured u 8.4e-6 / 1.4e-3, U_b 5.6e-7 / 5.5e-7 of the FP64 values
    assert e32 < 2.5e-5
    assert e16 < 4.5e-3
    assert b32 < 2e-6 and b16 < 2e-6


def test_ensemble_conservation_and_launch_forms(rr):
    """The 64-instance ensemble: batched = per-instance launches bit for bit
    (FP32), mass and momentum of every periodic box conserved, kinetic energy
    bounded."""
    from rrlbm import configs
    inst = configs.ensemble_instances()
    res = {}
    for mode in ("batched", "sequential"):
        with rr.Batch(inst, storage=1, keep_reference=False) as b:
            b.init_taylor_green_3d()
            d0 = [b.diag(k) for k in range(len(inst))]
            if mode == "batched":
                b.step(20)
            else:
                b.step_sequential(20)
            res[mode] = [b.diag(k) for k in range(len(inst))]
    for k, (lx, ly, lz, tau, u0) in enumerate(inst):
        a, s = res["batched"][k], res["sequential"][k]
        assert a["kinetic_energy"] == s["kinetic_energy"] and a["mass_excess"] == s["mass_excess"], k
        n = lx * ly * lz
        assert abs(a["mass_excess"] - d0[k]["mass_excess"]) < 1e-6 * n ** 0.5 + 1e-9, k
        assert np.abs(np.array(a["momentum"])).max() < 1e-7 * n * u0, k
        # uniform initial density: acoustic transients move a little energy
        # between kinetic and compressive parts over the first steps
        assert abs(a["kinetic_energy"] / d0[k]["kinetic_energy"] - 1) < 0.05, k
