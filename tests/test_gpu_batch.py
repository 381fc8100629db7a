"""Batched ensembles: one launch per step over many periodic D3Q27 boxes of
uneven size and relaxation time must give, bit for bit, what per-instance
launches of the same kernel give, and equal the oracle's reference-order
steps in FP64; the hipGraph replay must equal eager launches."""
import numpy as np
import pytest

import oracle
from helpers import maxabs

pytestmark = pytest.mark.gpu

# lx = 64 and 128: rows of whole wavefronts (row-scalar addressing, rows grouped in the
# default population layout); the others take per-lane addressing
INSTANCES = [(8, 8, 8, 0.51, 0.05), (12, 10, 9, 0.6, 0.03), (16, 16, 16, 0.55, 0.08), (7, 5, 6, 0.9, 0.02),
             (20, 12, 10, 0.52, 0.06), (64, 6, 5, 0.58, 0.04), (128, 50, 3, 0.53, 0.03)]


def fields_of(b, k, rr):
    s = b.instance(k)
    s.fused_unpack()
    return s.download(rr.Field.F), s.macroscopic()


# with a 96 x 64 x 40 box the work list has 1077 tiles: more than the 256 and
# 768 workgroups of persistent grids of 1 and 3 per compute unit on an MI355X
INSTANCES_BIG = INSTANCES + [(96, 64, 40, 0.56, 0.05)]


@pytest.mark.parametrize("storage", [0, 1, 2])
def test_batched_equals_sequential_and_graph(rr, storage):
    """FP16: the stochastic-rounding dither depends on (cell, step, instance)
    only, so the launch forms stay bit-identical; the graph of 2 steps is
    replayed three times before one eager step, so its replays must advance
    the dither."""
    steps = 7
    out = {}
    inst = INSTANCES_BIG
    for mode in ("batched", "sequential", "graph", "persistent1", "persistent3", "autotuned"):
        with rr.Batch(inst, storage=storage, keep_reference=True) as b:
            if mode.startswith("persistent"):
                b.set_grid(int(mode[-1]))   # 1 or 3 workgroups per CU walking the work list
            if mode == "autotuned":
                k, ms = b.autotune(trial_steps=2)
                assert k in (0, 2, 4, 8, 16) and ms > 0
            b.init_taylor_green_3d()
            if mode in ("batched", "persistent1", "persistent3", "autotuned"):
                b.step(steps)
            elif mode == "sequential":
                b.step_sequential(steps)
            else:
                b.graph_capture(2)
                b.graph_launch(3)
                b.step(1)
            out[mode] = [fields_of(b, k, rr) for k in range(len(inst))]
    for k in range(len(inst)):
        for mode in ("sequential", "graph", "persistent1", "persistent3", "autotuned"):
            assert np.array_equal(out["batched"][k][0], out[mode][k][0]), (k, mode)


def test_batched_fp64_matches_oracle(rr):
    steps = 9
    with rr.Batch(INSTANCES, storage=0, keep_reference=True) as b:
        b.init_taylor_green_3d()
        b.step(steps - 1)
        got = [fields_of(b, k, rr) for k in range(len(INSTANCES))]
    for k, (lx, ly, lz, tau, u0) in enumerate(INSTANCES):
        g = oracle.Fields(oracle.D3Q27_CUDA, lx, ly, lz)
        g.init_tg_fields(u0, three_d=True)
        g.init_equilibrium()
        g.step(steps, tau, walls=(0, 0, 0), wall_rr=0)
        assert maxabs(got[k][0], g.f) < 1e-13
        assert maxabs(got[k][1][1], g.ux) < 1e-13


def test_graph_replays_and_released_reference(rr):
    """Several replays of a captured graph (with eager work and synchronisation
    between them) match eager batched steps; released reference fields make
    field access an error but keep diagnostics working."""
    with rr.Batch(INSTANCES, storage=1, keep_reference=False) as b:
        b.init_taylor_green_3d()
        b.graph_capture(4)
        ke = []
        for rep in range(3):
            b.graph_launch(1)
            ke.append([b.diag(k)["kinetic_energy"] for k in range(len(INSTANCES))])
        with pytest.raises(rr.RRLBMError):
            b.instance(0).fused_unpack()
    with rr.Batch(INSTANCES, storage=1, keep_reference=False) as e:
        e.init_taylor_green_3d()
        ke_e = []
        for rep in range(3):
            e.step(4)
            ke_e.append([e.diag(k)["kinetic_energy"] for k in range(len(INSTANCES))])
    assert np.array_equal(np.array(ke), np.array(ke_e))
