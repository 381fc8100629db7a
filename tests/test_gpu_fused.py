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
    # rows of 64 cells: each wavefront is one row (scalar row addressing); the
    # cell counts leave the last workgroup partly empty
    "channel_q27_rows64": (oracle.D3Q27_CUDA, (64, 6, 5), dict(wall_y=1, wall_rr=1), (0, 1, 0), -1),
    "periodic_q27_rows64": (oracle.D3Q27_CUDA, (64, 5, 3), dict(wall_y=0, wall_rr=0), (0, 0, 0), -1),
    "cavity_q27_rows128": (oracle.D3Q27_CUDA, (128, 5, 3), dict(wall_x=1, wall_y=1, wall_z=1, wall_rr=1, lid_face=3,
                                                                lid_u=(0.05, 0, 0)), (1, 1, 1), 3),
    "cavity_q19_rows64": (oracle.D3Q19, (64, 7, 5), dict(wall_x=1, wall_y=1, wall_z=1, wall_rr=1, lid_face=3,
                                                         lid_u=(0.05, 0, 0)), (1, 1, 1), 3),
    "periodic_q19_rows64": (oracle.D3Q19, (64, 5, 3), dict(wall_y=0, wall_rr=0), (0, 0, 0), -1),
}


# population layouts: library default, structure of arrays without padding,
# groups of 3 rows with 48 B of padding (a partial last group in most cases)
LAYOUTS = {"default": {}, "soa": dict(pop_group_rows=-1, pop_pad_bytes=0),
           "rows3": dict(pop_group_rows=3, pop_pad_bytes=48)}


@pytest.mark.parametrize("layout", sorted(LAYOUTS))
@pytest.mark.parametrize("name", sorted(CASES))
def test_fused_fp64_matches_reference_order(rr, name, layout):
    kind, dims, params, walls, lid = CASES[name]
    params = dict(params, **LAYOUTS[layout])
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


@pytest.mark.parametrize("lagged,lx", [(0, 16), (1, 16), (0, 64), (1, 64)])
def test_fused_les(rr, lagged, lx):
    """WALE in the fused path: reference order (two launches) is exact; the
    one-launch lagged-gradient variant differs by the one-step lag only
    (lx = 64: rows of one wavefront)."""
    g = channel_state(oracle.D3Q27_CUDA, lx, 12, 8, u_tau=0.01, visc=0.002, seed=7, force=(2e-5, 0, 0))
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


@pytest.mark.parametrize("lx", [40, 64])
def test_fused_les_walls_on_every_axis(rr, lx):
    """WALE's one-sided derivatives at walls on all three axes (with rows of 64
    cells the x walls are per lane, the y and z walls uniform per wavefront):
    the fused FP64 step in the reference order (velocity pass, then the step)
    equals the drop-in kernels' step (LES_WALE, collision_RR, streaming, BB,
    macrovar, RRBC, macrovar_BN) to 1e-12.  The drop-in LES_WALE applies the
    reference's y-wall form (LES.cu:19-33, pinned against the oracle) to each
    walled axis."""
    g = random_state(oracle.D3Q27_CUDA, lx, 12, 10, seed=31, umax=0.04, force=(1e-5, 0, 0))
    kw = dict(lx=lx, ly=12, lz=10, npop=27, storage=0, tau=0.51, force=(1e-5, 0, 0), wall_x=1, wall_y=1, wall_z=1,
              wall_rr=1, lid_face=3, lid_u=(0.05, 0, 0), use_les=1, les_lagged=0)
    steps = 5
    with rr.Solver(**kw) as s:
        upload_state(s, g)
        s.step_reference(steps)
        ref_f, ref_m = s.download(rr.Field.F), s.macroscopic()
    with rr.Solver(**kw) as s:
        upload_state(s, g)
        s.fused_pack()
        s.fused_step(steps - 1)
        s.fused_unpack()
        f, m = s.download(rr.Field.F), s.macroscopic()
    ef = maxabs(f, ref_f)
    em = max(maxabs(a, b) for a, b in zip(m, ref_m))
    assert np.abs(ref_m[1]).max() > 1e-3  # a flow with gradients at every wall
    assert ef < 1e-12 and em < 1e-12, (ef, em)


@pytest.mark.parametrize("lx", [32, 64])
def test_fused_fp32_fp16_vs_fp64(rr, lx):
    """FP32 and FP16-shifted storage (FP32 arithmetic) against the FP64 run of
    the same kernels on a channel state, 50 steps."""
    g = channel_state(oracle.D3Q27_CUDA, lx, 24, 16, u_tau=0.006, visc=0.0016, seed=3, force=(1e-5, 0, 0))
    tau = 0.5 + 3 * 0.0016
    res = {}
    for st in (0, 1, 2):
        _, m, diag = run_fused(rr, g, 50, storage=st, tau=tau, force=(1e-5, 0, 0), wall_y=1, wall_rr=1)
        res[st] = (m, diag)
    umax = np.abs(res[0][0][1]).max()
    e32 = maxabs(res[1][0][1], res[0][0][1]) / umax
    e16 = maxabs(res[2][0][1], res[0][0][1]) / umax
    ke = res[0][1]["kinetic_energy"]
    k32 = abs(res[1][1]["kinetic_energy"] - ke) / ke
    k16 = abs(res[2][1]["kinetic_energy"] - ke) / ke
    print("FP32/FP16 vs FP64 (lx %d): u %.2e / %.2e, KE %.2e / %.2e" % (lx, e32, e16, k32, k16))
    # about 3x the measured errors on MI355X (u 1.4e-6 / 9.9e-4, KE 1.2e-6 / 1.6e-5)
    assert e32 < 5e-6, e32
    assert e16 < 3e-3, e16
    assert k32 < 4e-6, k32
    assert k16 < 5e-5, k16


@pytest.mark.parametrize("lagged", [0, 1])
def test_fused_les_fp32_fp16_vs_fp64(rr, lagged):
    """FP32 and FP16-shifted storage with WALE (lagged and reference order)
    against the FP64 run of the same kernels on a channel state, 50 steps."""
    g = channel_state(oracle.D3Q27_CUDA, 64, 24, 16, u_tau=0.006, visc=0.0016, seed=3, force=(1e-5, 0, 0))
    tau = 0.5 + 3 * 0.0016
    res = {}
    for st in (0, 1, 2):
        _, m, diag = run_fused(rr, g, 50, storage=st, tau=tau, force=(1e-5, 0, 0), wall_y=1, wall_rr=1, use_les=1,
                               les_lagged=lagged)
        res[st] = (m, diag)
    umax = np.abs(res[0][0][1]).max()
    e32 = maxabs(res[1][0][1], res[0][0][1]) / umax
    e16 = maxabs(res[2][0][1], res[0][0][1]) / umax
    ke = res[0][1]["kinetic_energy"]
    k32 = abs(res[1][1]["kinetic_energy"] - ke) / ke
    k16 = abs(res[2][1]["kinetic_energy"] - ke) / ke
    print("LES (lagged %d) FP32/FP16 vs FP64: u %.2e / %.2e, KE %.2e / %.2e" % (lagged, e32, e16, k32, k16))
    # about 3x the measured errors (u 1.4e-6 / 8.6e-4, KE 1.2e-6 / 1.2e-5)
    assert e32 < 5e-6 and e16 < 3e-3 and k32 < 4e-6 and k16 < 4e-5, (e32, e16, k32, k16)


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


@pytest.mark.parametrize("lx", [20, 64])
@pytest.mark.parametrize("storage", [1, 2])
@pytest.mark.parametrize("les", [0, 1])
def test_population_layouts_are_bit_identical(rr, lx, storage, les):
    """The population layout only moves bytes: structure of arrays, grouped
    rows with a partial last group and padded groups give bit-identical
    states (FP32 and FP16 storage, walls, LES, per-lane and row-scalar
    addressing)."""
    g = channel_state(oracle.D3Q27_CUDA, lx, 12, 7, u_tau=0.01, visc=0.002, seed=5, force=(2e-5, 0, 0))
    tau = 0.5 + 3 * 0.002
    out = []
    for lay in (dict(pop_group_rows=-1, pop_pad_bytes=0), dict(pop_group_rows=5, pop_pad_bytes=0),
                dict(pop_group_rows=2, pop_pad_bytes=256), {}):
        f, m, _ = run_fused(rr, g, 15, storage=storage, tau=tau, force=(2e-5, 0, 0), wall_y=1, wall_rr=1,
                            use_les=les, **lay)
        out.append((f, m[1]))
    for f, u in out[1:]:
        assert np.array_equal(f, out[0][0])
        assert np.array_equal(u, out[0][1])


def test_layout_parameters_are_reported_and_checked(rr):
    with rr.Solver(lx=128, ly=64, lz=8, storage=1) as s:
        p = s.params
        assert p.pop_group_rows == 48 and p.pop_pad_bytes == 0   # 48 rows of 128 = 6144 cells
    with rr.Solver(lx=100, ly=10, lz=4, storage=2) as s:
        p = s.params
        assert p.pop_group_rows == -1 and p.pop_pad_bytes == 2048
    for bad in (dict(pop_group_rows=-2), dict(pop_pad_bytes=24), dict(pop_pad_bytes=2 << 20), dict(xcd_remap=-2),
                dict(xcd_remap=65537), dict(cells_per_thread=3), dict(persistent_wg_per_cu=65)):
        with pytest.raises(rr.RRLBMError) as ei:
            rr.Solver(lx=8, ly=8, lz=8, **bad)
        assert ei.value.status == 1


# 128 x 64 x 40 = 327 680 cells = 1280 blocks of 256 (640 of the two-cell
# kernel): more blocks than a persistent grid of 1 or 3 workgroups per compute
# unit has workgroups on the 256 CUs of an MI355X, so those grids stride.
PERSIST_DIMS = (128, 64, 40)


@pytest.mark.parametrize("shape", [dict(persistent_wg_per_cu=1), dict(persistent_wg_per_cu=3, xcd_remap=1),
                                   dict(cells_per_thread=2, persistent_wg_per_cu=1), dict(xcd_remap=1),
                                   dict(xcd_remap=4), dict(xcd_remap=16, persistent_wg_per_cu=1),
                                   dict(xcd_remap=3), dict(xcd_remap=24, persistent_wg_per_cu=1),
                                   dict(cells_per_thread=2, xcd_remap=2)])
def test_launch_shapes_are_bit_identical(rr, shape):
    """Persistent grids (workgroups striding over the blocks; the unpack and
    diagnostics passes keep one workgroup per block), the XCD-remapped block
    order and two cells per thread compute the same state as one workgroup
    per block."""
    g = channel_state(oracle.D3Q27_CUDA, *PERSIST_DIMS, u_tau=0.01, visc=0.002, seed=9, force=(2e-5, 0, 0))
    tau = 0.5 + 3 * 0.002
    kw = dict(storage=1, tau=tau, force=(2e-5, 0, 0), wall_y=1, wall_rr=1)
    f0, m0, d0 = run_fused(rr, g, 12, **kw)
    f1, m1, d1 = run_fused(rr, g, 12, **kw, **shape)
    if shape.get("cells_per_thread") == 2:
        # the two-cell kernel is compiled separately: its FMA contractions may
        # round differently (FP32: a few ulp)
        assert maxabs(f0, f1) < 1e-6 and maxabs(m0[1], m1[1]) < 1e-6
        return
    assert np.array_equal(f0, f1) and np.array_equal(m0[1], m1[1])
    assert d0["kinetic_energy"] == d1["kinetic_energy"] and d0["mass_excess"] == d1["mass_excess"]


@pytest.mark.parametrize("les_lagged", [0, 1])
def test_persistent_grids_cover_every_pass(rr, les_lagged):
    """With LES in the reference order (a velocity pass, then the step) or
    lagged, persistent grids of 1 and 4 workgroups per compute unit on a grid
    of more blocks than that (1280 against 256 and 1024 workgroups) give, bit
    for bit, the populations, fields and diagnostics of one workgroup per
    block: every pass covers every block."""
    g = channel_state(oracle.D3Q27_CUDA, *PERSIST_DIMS, u_tau=0.01, visc=0.002, seed=19, force=(2e-5, 0, 0))
    kw = dict(storage=1, tau=0.5 + 3 * 0.002, force=(2e-5, 0, 0), wall_y=1, wall_rr=1, use_les=1,
              les_lagged=les_lagged)
    f0, m0, d0 = run_fused(rr, g, 6, persistent_wg_per_cu=0, **kw)
    for k in (1, 4):
        f1, m1, d1 = run_fused(rr, g, 6, persistent_wg_per_cu=k, **kw)
        assert np.array_equal(f0, f1), k
        for a, b in zip(m0, m1):
            assert np.array_equal(a, b), k
        assert d0 == d1, k


def test_autotune_picks_a_shape_with_identical_results(rr):
    """rrlbm_autotune times layouts and launch shapes on the context's own
    buffers before packing, keeps the fastest (identical arithmetic, so an
    identical state), caches the choice, and refuses to run after packing."""
    g = channel_state(oracle.D3Q27_CUDA, 128, 16, 8, u_tau=0.01, visc=0.002, seed=4, force=(2e-5, 0, 0))
    tau = 0.5 + 3 * 0.002
    kw = dict(storage=2, tau=tau, force=(2e-5, 0, 0), wall_y=1, wall_rr=1)
    ref_f, ref_m, _ = run_fused(rr, g, 10, **kw)
    with rr.Solver(lx=g.lx, ly=g.ly, lz=g.lz, **kw) as s:
        t = s.autotune(trial_steps=2)
        assert t["candidates"] > 10 and t["ms_per_step"] > 0
        p = s.params
        assert (p.pop_group_rows, p.pop_pad_bytes, p.cells_per_thread, p.xcd_remap, p.persistent_wg_per_cu) == (
            t["pop_group_rows"], t["pop_pad_bytes"], t["cells_per_thread"], t["xcd_remap"], t["persistent_wg_per_cu"])
        upload_state(s, g)
        s.fused_pack()
        s.fused_step(9)
        s.fused_unpack()
        if t["cells_per_thread"] == 1:
            assert np.array_equal(s.download(rr.Field.F), ref_f)
        else:  # separately compiled two-cell kernel: rounding may differ (FP16 storage)
            assert maxabs(s.download(rr.Field.F), ref_f) < 1e-3
        with pytest.raises(rr.RRLBMError) as ei:
            s.autotune(trial_steps=2)
        assert ei.value.status == 3
    with rr.Solver(lx=g.lx, ly=g.ly, lz=g.lz, **kw) as s:
        t2 = s.autotune(trial_steps=2)
        assert t2["candidates"] == 0 and t2["pop_group_rows"] == t["pop_group_rows"]


@pytest.mark.parametrize("storage", [1, 2])
@pytest.mark.parametrize("les", [0, 1])
def test_graph_replays_match_eager_steps(rr, les, storage):
    """Fused steps captured into a HIP graph: three replays of 4 steps, with
    eager work and synchronisation between them, give the state of 12 eager
    steps bit for bit (FP32, and FP16 with stochastic rounding, whose replays
    take new dither from the device step counter; with and without the lagged
    LES velocity buffers); an odd capture length refuses its second replay."""
    g = channel_state(oracle.D3Q27_CUDA, 64, 12, 8, u_tau=0.01, visc=0.002, seed=21, force=(2e-5, 0, 0))
    kw = dict(storage=storage, tau=0.5 + 3 * 0.002, force=(2e-5, 0, 0), wall_y=1, wall_rr=1, use_les=les)
    f_eager, m_eager, d_eager = run_fused(rr, g, 13, **kw)
    with rr.Solver(lx=g.lx, ly=g.ly, lz=g.lz, **kw) as s:
        upload_state(s, g)
        s.fused_pack()
        s.graph_capture(4)
        ke = []
        for rep in range(3):
            s.graph_launch(1)
            ke.append(s.fused_diag()["kinetic_energy"])
        s.fused_unpack()
        assert np.array_equal(s.download(rr.Field.F), f_eager)
        assert ke[-1] == d_eager["kinetic_energy"]
        s.graph_capture(3)
        s.graph_launch(1)
        with pytest.raises(rr.RRLBMError) as ei:
            s.graph_launch(1)
        assert ei.value.status == 3
        if storage == 1 and not les:
            # a replay completes a step collided under the uniform force: refused
            # after forcing_GP steps until a plain step has run
            s.graph_capture(2)
            s.fused_step_gp(0, 0.3)
            s.fused_step_gp(1, 0.3)
            with pytest.raises(rr.RRLBMError) as ei:
                s.graph_launch(1)
            assert ei.value.status == 3
            s.fused_step(2)
            s.graph_launch(1)
            # a new population layout (autotuning) drops the captured graph
            s.init_rest()
            s.autotune(trial_steps=1)
            s.fused_pack()
            with pytest.raises(rr.RRLBMError) as ei:
                s.graph_launch(1)
            assert ei.value.status == 3


@pytest.mark.parametrize("force", [(2e-5, 1e-6, 0), (2e-5, 0, 0)])
@pytest.mark.parametrize("lx", [12, 64])
def test_fused_fdbc_matches_reference_order(rr, lx, force):
    """FDBC at the y walls (wall_rr = 2) in the fused path: a row pass for the
    velocities of rows 1, 2, LY-2, LY-3, then the step with the wall rows
    reconstructed; against the oracle's reference-order step (force along x
    only: the step kernel without the y and z force terms)."""
    g = channel_state(oracle.D3Q27_CUDA, lx, 10, 6, u_tau=0.01, visc=0.002, seed=9, force=force)
    tau = 0.5 + 3 * 0.002
    steps = 8
    got_f, got_m, diag = run_fused(rr, g, steps, storage=0, tau=tau, force=force, wall_y=1, wall_rr=2)
    g.step(steps, tau, walls=(0, 1, 0), wall_rr=2)
    assert maxabs(got_f, g.f) < 1e-12
    for a, b in zip(got_m, (g.rho, g.ux, g.uy, g.uz)):
        assert maxabs(a, b) < 1e-12
    assert abs(diag["mass_excess"] - g.diag()[4]) < 1e-10


def test_fused_fdbc_fp32_fp16_vs_fp64(rr):
    g = channel_state(oracle.D3Q27_CUDA, 64, 24, 16, u_tau=0.006, visc=0.0016, seed=3, force=(1e-5, 0, 0))
    tau = 0.5 + 3 * 0.0016
    res = {}
    for st in (0, 1, 2):
        _, m, _ = run_fused(rr, g, 50, storage=st, tau=tau, force=(1e-5, 0, 0), wall_y=1, wall_rr=2)
        res[st] = m
    umax = np.abs(res[0][1]).max()
    e32, e16 = maxabs(res[1][1], res[0][1]) / umax, maxabs(res[2][1], res[0][1]) / umax
    print("FDBC FP32/FP16 vs FP64: u %.2e / %.2e" % (e32, e16))
    assert e32 < 5e-6      # measured 1.3e-6
    assert e16 < 3e-3      # measured 8.7e-4


@pytest.mark.parametrize("dims", [(10, 9, 7), (64, 5, 3)])
def test_fused_d3q19_without_force(rr, dims):
    """D3Q19 without body force runs the moment-space form (rr_post_q19):
    periodic box against the oracle's reference-order step, FP64."""
    g = random_state(oracle.D3Q19, *dims, seed=13, force=(0.0, 0.0, 0.0))
    tau, steps = 0.55, 12
    got_f, got_m, _ = run_fused(rr, g, steps, storage=0, tau=tau, force=(0.0, 0.0, 0.0), wall_y=0, wall_rr=0)
    g.step(steps, tau, walls=(0, 0, 0), wall_rr=0)
    assert maxabs(got_f, g.f) < 1e-12
    for a, b in zip(got_m, (g.rho, g.ux, g.uy, g.uz)):
        assert maxabs(a, b) < 1e-12


@pytest.mark.parametrize("les", [False, True])
def test_fused_forcing_gp_matches_reference_order(rr, les):
    """forcing_GP computed inside the fused kernel (rrlbm_fused_step_gp)
    against the drop-in loop of main.cu:33-84 (forcing_GP fields + the
    reference kernels): GP steps with a phase change, then the uniform force."""
    lx, ly, lz = 32, 24, 8
    g = channel_state(oracle.D3Q27_CUDA, lx, ly, lz, u_tau=0.01, visc=0.002, seed=5, force=(2e-5, 0, 0))
    tau = 0.5 + 3 * 0.002
    kw = dict(lx=lx, ly=ly, lz=lz, storage=0, tau=tau, force=(2e-5, 0, 0), wall_y=1, wall_rr=1, use_les=int(les),
              les_lagged=0)
    schedule = [(istep, 0.31 if istep < 607 else 0.77) for istep in range(601, 613)]   # phase change at 607
    gp_end = 610
    with rr.Solver(**kw) as s:
        upload_state(s, g)
        for istep, ph in schedule:
            if istep < gp_end:
                s.forcing_GP(istep, ph)
            else:
                s.forcing()
            s.step_reference(1)
        ref_f = s.download(rr.Field.F)
        ref_m = s.macroscopic()
    with rr.Solver(**kw) as s:
        upload_state(s, g)
        istep, ph = schedule[0]
        s.forcing_GP(istep, ph)
        s.fused_pack()
        for istep, ph in schedule[1:]:
            if istep < gp_end:
                s.fused_step_gp(istep, ph)
            else:
                s.fused_step(1)
        s.fused_unpack()
        f = s.download(rr.Field.F)
        m = s.macroscopic()
    assert maxabs(f, ref_f) < 1e-12
    for a, b in zip(m, ref_m):
        assert maxabs(a, b) < 1e-12
    # the force mattered: a run without it differs
    with rr.Solver(**kw) as s:
        upload_state(s, g)
        s.fused_pack()
        s.fused_step(len(schedule) - 1)
        s.fused_unpack()
        assert maxabs(s.macroscopic()[1], ref_m[1]) > 1e-7


@pytest.mark.parametrize("dims,walls", [((1, 1, 1), (0, 0, 0)), ((1, 6, 1), (0, 0, 0)), ((3, 1, 2), (0, 0, 0)),
                                        ((1, 4, 1), (0, 1, 0)), ((2, 2, 2), (1, 1, 1)), ((65, 3, 1), (0, 1, 0))])
@pytest.mark.parametrize("kind", [oracle.D3Q27_CUDA, oracle.D3Q19])
def test_fused_degenerate_grids(rr, dims, walls, kind):
    """Edge sizes: single cells and single rows (periodic images of the cell
    itself), two-cell walled boxes, a row of 65 (one full wavefront plus one
    cell): the fused FP64 step against the oracle's reference-order step."""
    force = (0.0, 0.0, 0.0) if kind == oracle.D3Q19 else (2e-5, -1e-5, 3e-6)
    g = random_state(kind, *dims, seed=3, force=force)
    wall_rr = 1 if any(walls) else 0
    got_f, got_m, _ = run_fused(rr, g, 6, storage=0, tau=0.6, force=force, wall_x=walls[0], wall_y=walls[1],
                                wall_z=walls[2], wall_rr=wall_rr)
    g.step(6, 0.6, walls=walls, wall_rr=wall_rr)
    assert maxabs(got_f, g.f) < 1e-12
    for a, b in zip(got_m, (g.rho, g.ux, g.uy, g.uz)):
        assert maxabs(a, b) < 1e-12
