"""Two fused steps per launch (rrlbm_params.steps_per_pass = 2).

The two-step kernel marches through z over tiles of 64 x 4 cells, computing
the first step on the tile and a one-cell halo into a ring of planes in LDS
and the second step from that ring.  It calls the one-step kernel's per-cell
functions, so its states must equal those of one-step launches: bit for bit
where both kernels contract the same multiply-adds, else to an ulp (FP32 and
FP64 storage, walls on any axis with the moving lid, uniform force
along x or in any direction, odd step counts that end with a one-step launch,
z chunks of 1, 2, 4, 64 planes and the default, grids of one and two planes,
several tiles in x and y, one tile in y), for the RR
collision and for the BGK and sixth-order RR operators.  The FP64 RR run is
also checked against the oracle's reference-order step (1e-12)."""
import numpy as np
import pytest

import oracle
from helpers import channel_state, maxabs, random_state, upload_state

pytestmark = pytest.mark.gpu

CASES = {
    # name: (dims, params, force, oracle walls, lid)
    "channel": ((64, 8, 6), dict(wall_y=1, wall_rr=1), (3e-5, 0.0, 0.0), (0, 1, 0), -1),
    "channel_force_xyz": ((64, 8, 5), dict(wall_y=1, wall_rr=1), (3e-5, 1e-6, -2e-6), (0, 1, 0), -1),
    "channel_bb_only": ((128, 12, 3), dict(wall_y=1, wall_rr=0), (3e-5, 0.0, 0.0), (0, 1, 0), -1),
    "periodic": ((128, 4, 7), dict(wall_y=0, wall_rr=0), (0.0, 0.0, 0.0), (0, 0, 0), -1),
    "periodic_one_plane": ((64, 4, 1), dict(wall_y=0, wall_rr=0), (2e-5, 0.0, 0.0), (0, 0, 0), -1),
    "cavity": ((64, 8, 6), dict(wall_x=1, wall_y=1, wall_z=1, wall_rr=1, lid_face=3, lid_u=(0.05, 0.0, 0.0)),
               (0.0, 0.0, 0.0), (1, 1, 1), 3),
    # three tiles in x and y (nine work items, fewer than the eight XCDs' runs
    # fill evenly), walls on x and y
    "duct_3x3_tiles": ((192, 12, 5), dict(wall_x=1, wall_y=1, wall_rr=1), (3e-5, 0.0, 0.0), (1, 1, 0), -1),
    # one tile in y: both halo rows wrap onto the wall rows
    "channel_one_tile_row": ((64, 4, 5), dict(wall_y=1, wall_rr=1), (3e-5, 0.0, 0.0), (0, 1, 0), -1),
    # two planes between z walls: every chunk border is a wall plane
    "two_planes_z_walls": ((64, 8, 2), dict(wall_y=0, wall_z=1, wall_rr=1), (0.0, 0.0, 0.0), (0, 0, 1), -1),
}


def run(rr, g, steps, storage, **params):
    with rr.Solver(lx=g.lx, ly=g.ly, lz=g.lz, npop=27, storage=storage, tau=0.53, **params) as s:
        upload_state(s, g)
        s.fused_pack()
        s.fused_step(steps - 1)
        s.fused_unpack()
        return s.download(rr.Field.F), s.macroscopic()


def state(name):
    dims, params, force, walls, lid = CASES[name]
    return random_state(oracle.D3Q27_CUDA, *dims, seed=5, force=force), dims, params, force, walls, lid


@pytest.mark.parametrize("planes", [0, 1, 2, 4, 64])
@pytest.mark.parametrize("storage", [0, 1])
@pytest.mark.parametrize("name", sorted(CASES))
def test_two_step_equals_one_step_launches(rr, name, storage, planes):
    g, dims, params, force, walls, lid = state(name)
    steps = 8  # pack + 7 fused steps: three two-step launches and one one-step launch
    f1, m1 = run(rr, g, steps, storage, force=force, **params)
    f2, m2 = run(rr, g, steps, storage, force=force, steps_per_pass=2, pass_planes=planes, **params)
    ef = maxabs(f1, f2)
    em = max(maxabs(a, b) for a, b in zip(m1, m2))
    print("%s storage %d planes %d: max |df| %.3g, max |d macro| %.3g, identical %s"
          % (name, storage, planes, ef, em, np.array_equal(f1, f2)))
    tol = 1e-15 if storage == 0 else 1e-6
    assert ef < tol and em < tol, (ef, em)


OP_CASES = {
    # name: (dims, params, force); BGK walls are plain bounce-back (wall_rr 0)
    "bgk_channel": ((64, 8, 6), dict(wall_y=1, wall_rr=0, collision_model=1), (3e-5, 0.0, 0.0)),
    "bgk_periodic_force_xyz": ((128, 4, 5), dict(wall_rr=0, collision_model=1), (3e-5, 1e-6, -2e-6)),
    "rr6_channel": ((64, 8, 6), dict(wall_y=1, wall_rr=1, rr_order=6), (3e-5, 0.0, 0.0)),
    "rr6_channel_force_xyz": ((64, 8, 5), dict(wall_y=1, wall_rr=1, rr_order=6), (3e-5, 1e-6, -2e-6)),
    "rr6_cavity": ((64, 8, 6), dict(wall_x=1, wall_y=1, wall_z=1, wall_rr=1, lid_face=3, lid_u=(0.05, 0.0, 0.0),
                                    rr_order=6), (0.0, 0.0, 0.0)),
}


@pytest.mark.parametrize("planes", [0, 2])
@pytest.mark.parametrize("storage", [0, 1])
@pytest.mark.parametrize("name", sorted(OP_CASES))
def test_two_step_operators_equal_one_step_launches(rr, name, storage, planes):
    """The BGK operator and the sixth-order RR collision (with its wall
    regularisation) on the two-step path equal one-step launches of the same
    operator, whose kernels are checked against the drop-in reference order."""
    dims, params, force = OP_CASES[name]
    g = random_state(oracle.D3Q27_CUDA, *dims, seed=6, force=force)
    steps = 8
    f1, m1 = run(rr, g, steps, storage, force=force, **params)
    f2, m2 = run(rr, g, steps, storage, force=force, steps_per_pass=2, pass_planes=planes, **params)
    ef = maxabs(f1, f2)
    em = max(maxabs(a, b) for a, b in zip(m1, m2))
    print("%s storage %d planes %d: max |df| %.3g, max |d macro| %.3g, identical %s"
          % (name, storage, planes, ef, em, np.array_equal(f1, f2)))
    tol = 1e-15 if storage == 0 else 1e-6
    assert ef < tol and em < tol, (ef, em)


@pytest.mark.parametrize("name", ["channel", "cavity", "periodic", "duct_3x3_tiles", "two_planes_z_walls"])
def test_two_step_fp64_matches_oracle(rr, name):
    g, dims, params, force, walls, lid = state(name)
    steps = 9
    f, m = run(rr, g, steps, 0, force=force, steps_per_pass=2, **params)
    g.step(steps, 0.53, walls=walls, wall_rr=params.get("wall_rr", 0), lid_face=lid,
           lid_u=params.get("lid_u", (0.0, 0.0, 0.0)))
    assert maxabs(f, g.f) < 1e-12
    for a, b in zip(m, (g.rho, g.ux, g.uy, g.uz)):
        assert maxabs(a, b) < 1e-12


@pytest.mark.parametrize("nsteps,times", [(4, 3), (2, 1), (6, 1), (5, 1)])
def test_two_step_graph_replays_equal_eager_steps(rr, nsteps, times):
    """A captured graph of nsteps steps (two-step launches, and a one-step
    launch for an odd count), replayed `times` times, equals as many eager
    steps followed by one more; a graph whose launches swap the buffers an odd
    number of times (2 or 6 steps: 1 or 3 launches; 5 steps: 3) refuses a
    second replay."""
    g, dims, params, force, walls, lid = state("channel")
    kw = dict(lx=g.lx, ly=g.ly, lz=g.lz, npop=27, storage=1, tau=0.53, force=force, steps_per_pass=2, **params)
    out = []
    for graph in (False, True):
        with rr.Solver(**kw) as s:
            upload_state(s, g)
            s.fused_pack()
            if graph:
                s.graph_capture(nsteps)
                s.graph_launch(times)
                if nsteps % 4 != 0:
                    with pytest.raises(rr.RRLBMError) as ei:
                        s.graph_launch(1)
                    assert ei.value.status == 3
            else:
                s.fused_step(nsteps * times)
            s.fused_step(1)
            s.fused_unpack()
            out.append(s.download(rr.Field.F))
    assert np.array_equal(out[0], out[1]), maxabs(out[0], out[1])


def test_autotune_times_two_step_launches(rr):
    """The autotuner's candidates include two steps per launch (FP32 channel);
    the context takes the reported choice, and its states equal one-step
    launches of the default shape."""
    g = channel_state(oracle.D3Q27_CUDA, 128, 16, 24, u_tau=0.01, visc=0.002, seed=4, force=(2e-5, 0, 0))
    tau = 0.5 + 3 * 0.002
    kw = dict(lx=g.lx, ly=g.ly, lz=g.lz, npop=27, storage=1, tau=tau, force=(2e-5, 0, 0), wall_y=1, wall_rr=1)
    out = []
    for tune in (False, True):
        with rr.Solver(**kw) as s:
            if tune:
                t = s.autotune(trial_steps=2)
                p = s.params
                assert (p.steps_per_pass, p.pass_planes) == (t["steps_per_pass"], t["pass_planes"])
                assert t["steps_per_pass"] in (1, 2)
            upload_state(s, g)
            s.fused_pack()
            s.fused_step(9)
            s.fused_unpack()
            out.append(s.download(rr.Field.F))
    assert maxabs(out[0], out[1]) < 1e-6


def test_two_step_after_forcing_gp_steps(rr):
    """forcing_GP steps run as one-step launches; the plain steps after them
    (one single step to complete the GP collision, then pairs) equal one-step
    launches of the same sequence."""
    g, dims, params, force, walls, lid = state("channel")
    out = []
    for spp in (1, 2):
        with rr.Solver(lx=g.lx, ly=g.ly, lz=g.lz, npop=27, storage=1, tau=0.53, force=force, steps_per_pass=spp,
                       **params) as s:
            upload_state(s, g)
            s.fused_pack()
            s.fused_step(3)
            s.fused_step_gp(40, 0.3)
            s.fused_step_gp(41, 0.3)
            s.fused_step(5)
            s.fused_unpack()
            out.append(s.download(rr.Field.F))
    assert np.array_equal(out[0], out[1]), maxabs(out[0], out[1])
