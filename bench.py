"""Throughput of the fused RR-LBM step on one MI355X.

Workload: turbulent channel DNS, D3Q27 RR, Re_tau = 180, 384 x 128 x 192
(x streamwise periodic, y half-way bounce-back walls with RR regularisation,
z periodic), u_tau = 0.0045, h = 64, nu = u_tau h / 180, uniform body force
u_tau^2/h, seeded log-law + streak + noise initial field, FP32 shifted
storage (the headline), with the FP16-scaled and FP64 storage variants timed
beside it.

One step = one fused launch over the whole grid (pull-stream, bounce-back,
macrovar, RRBC, macrovar_BN, RR collision).  value = MLUPS = cells * steps /
seconds of the timed region, the populations resident in HBM throughout.

Roofline: algorithmic bytes per lattice update B = 2 * 27 * sizeof(storage)
(each population read once and written once); achieved = B * cells / (mean
fused-kernel duration from HIP events on the library's stream); peak = 8 TB/s.
traffic = HBM bytes per launch measured by rocprofv3 (profiles/), when present.

CPU baseline: the oracle's reference-order step (C, OpenMP) on the same grid
for a bounded number of steps on the host cores.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "recursively-regularized-lbm_amd"))

LX, LY, LZ = 384, 128, 192
U_TAU = 0.0045
H = LY / 2
VISC = U_TAU * H / 180.0
TAU = 3.0 * VISC + 0.5
FORCE = U_TAU * U_TAU / H
SEED = 2024
PEAK_GBS = 8000.0
STORAGE_BYTES = {0: 8, 1: 4, 2: 2}
STORAGE_NAME = {0: "f64", 1: "f32", 2: "f16"}


def run_variant(rrlbm, storage, steps, warmup):
    s = rrlbm.Solver(lx=LX, ly=LY, lz=LZ, npop=27, storage=storage, tau=TAU, force=(FORCE, 0.0, 0.0),
                     wall_y=1, wall_rr=1, use_les=0)
    try:
        s.init_channel(U_TAU, VISC, SEED, 0.1, 0.01)
        s.fused_pack()
        s.fused_step(warmup)
        s.synchronize()
        t0 = time.perf_counter()
        kernel_ms = s.fused_step_timed(steps)
        s.synchronize()
        wall = time.perf_counter() - t0
        d = s.fused_diag()
    finally:
        s.close()
    n = LX * LY * LZ
    mlups = n * steps / wall / 1e6
    per_launch = kernel_ms / 1e3 / steps
    blu = 2 * 27 * STORAGE_BYTES[storage]
    achieved = blu * n / per_launch / 1e9
    return {
        "mlups": mlups,
        "ms_per_step": wall * 1e3 / steps,
        "kernel_ms": kernel_ms / steps,
        "bytes_per_lu": blu,
        "achieved_gbs": achieved,
        "frac": achieved / PEAK_GBS,
        "mass_excess": d["mass_excess"],
        "ubulk_over_utau": d["u_sum"][0] / n / U_TAU,
    }


def traffic_from_profiles(storage):
    """HBM bytes per launch from the committed rocprofv3 counter summary, if any."""
    path = os.path.join(REPO, "profiles", "fused_channel_384x128x192_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        t = json.load(fh)
    return t.get(STORAGE_NAME[storage], {}).get("hbm_bytes_per_launch")


def cpu_baseline(steps=2):
    """Oracle (C FP64, OpenMP) reference-order steps on the full grid."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle

    g = oracle.Fields(oracle.D3Q27_CUDA, LX, LY, LZ)
    g.init_channel_fields(U_TAU, VISC, SEED, 0.1, 0.01)
    g.init_equilibrium()
    g.set_force((FORCE, 0.0, 0.0))
    t0 = time.perf_counter()
    g.step(steps, TAU, walls=(0, 1, 0), wall_rr=1)
    dt = time.perf_counter() - t0
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": LX * LY * LZ * steps / dt / 1e6, "unit": "MLUPS", "cores": cores, "kind": "port",
            "sample": "%d reference-order steps (FP64 oracle, C/OpenMP) of the full %dx%dx%d grid, %.1f s"
                      % (steps, LX, LY, LZ, dt)}


def ensemble_instances(seed=5, count=64):
    """64 periodic 3D Taylor-Green boxes: N in {32,48,64,96,128,192}, Ma in
    U[0.01, 0.1] (u0 = Ma cs), Re in U[100, 1600], nu = u0 (N / 2 pi) / Re."""
    import numpy as np

    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = int(rng.choice([32, 48, 64, 96, 128, 192]))
        ma = rng.uniform(0.01, 0.1)
        re = rng.uniform(100.0, 1600.0)
        u0 = ma / np.sqrt(3.0)
        nu = u0 * (n / (2 * np.pi)) / re
        out.append((n, n, n, 3.0 * nu + 0.5, u0))
    return out


def run_ensemble(rrlbm, steps, warmup, storage=1):
    inst = ensemble_instances()
    res = {}
    with rrlbm.Batch(inst, storage=storage, keep_reference=False) as b:
        cells = b.cells()
        b.init_taylor_green_3d()
        b.step(warmup)
        b.step_sequential(warmup)
        res["batched_ms"] = b.time(0, steps)
        res["sequential_ms"] = b.time(1, steps)
        b.graph_capture(steps)
        b.graph_launch(1)
        res["graph_ms"] = b.time(2, 1)
        ke = [b.diag(k)["kinetic_energy"] for k in range(len(inst))]
    mlups = {k.replace("_ms", ""): cells * steps / (v / 1e3) / 1e6 for k, v in res.items()}
    return cells, mlups, ke


def main_ensemble(args):
    import rrlbm

    rrlbm.load()
    cells, m, ke = run_ensemble(rrlbm, args.steps, args.warmup)
    blu = 2 * 27 * 4
    best = max(m["graph"], m["batched"])
    line = {
        "metric": "MLUPS, batched 3D Taylor-Green ensemble (64 D3Q27 RR instances, FP32)",
        "value": round(best, 1), "unit": "MLUPS", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(cells / (best * 1e6) * 1e3, 5), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded ensemble)",
        "config": {"workload": "64 periodic D3Q27 TG instances, N in {32..192}, %d cells" % cells},
        "roofline": {"bound": "hbm", "achieved": round(best * 1e6 * blu / 1e9, 1), "peak": PEAK_GBS, "unit": "GB/s",
                     "frac": round(best * 1e6 * blu / 1e9 / PEAK_GBS, 4), "traffic": None, "bytes_per_lu": blu},
        "modes_mlups": {k: round(v, 1) for k, v in m.items()},
        "speedup_vs_sequential": round(best / m["sequential"], 3),
        "finite": bool(all(x == x for x in ke)),
    }
    print(json.dumps(line))


def main_copy():
    """Calibration workload for rocprofv3's FETCH_SIZE / WRITE_SIZE counters: a
    plain device-to-device copy of 1 GiB, five times."""
    import torch

    n = (1 << 30) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"workload": "copy", "bytes_read_per_copy": 4 * n, "bytes_written_per_copy": 4 * n, "copies": 5}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--variants", default="1,2,0", help="storage variants; the first is the headline")
    ap.add_argument("--workload", default="channel", choices=["channel", "ensemble", "copy"])
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "ensemble":
        return main_ensemble(args)
    if args.workload == "copy":
        return main_copy()
    import rrlbm

    rrlbm.load()
    variants = [int(v) for v in args.variants.split(",")]
    res = {st: run_variant(rrlbm, st, args.steps, args.warmup) for st in variants}
    head = res[variants[0]]
    st0 = variants[0]
    line = {
        "metric": "MLUPS + % HBM roofline, D3Q27 RR Re_tau=180 channel; FP32 vs FP16-shifted",
        "value": round(head["mlups"], 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms_per_step"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": STORAGE_NAME[st0],
        "data": "synthetic (seeded log-law + streaks + noise channel field)",
        "config": {"workload": "channel DNS Re_tau=180 D3Q27 RR, %dx%dx%d, %s shifted storage"
                               % (LX, LY, LZ, STORAGE_NAME[st0].upper()),
                   "lx": LX, "ly": LY, "lz": LZ, "npop": 27, "storage": STORAGE_NAME[st0], "u_tau": U_TAU,
                   "tau": TAU, "force_x": FORCE, "les": 0},
        "roofline": {"bound": "hbm", "achieved": round(head["achieved_gbs"], 1), "peak": PEAK_GBS, "unit": "GB/s",
                     "frac": round(head["frac"], 4), "traffic": traffic_from_profiles(st0),
                     "bytes_per_lu": head["bytes_per_lu"], "kernel_ms": round(head["kernel_ms"], 5)},
        "variants": {STORAGE_NAME[st]: {"mlups": round(r["mlups"], 1), "frac": round(r["frac"], 4),
                                         "achieved_gbs": round(r["achieved_gbs"], 1),
                                         "bytes_per_lu": r["bytes_per_lu"], "kernel_ms": round(r["kernel_ms"], 5),
                                         "ubulk_over_utau": round(r["ubulk_over_utau"], 4)}
                     for st, r in res.items()},
        "cpu_baseline": None if args.no_cpu else cpu_baseline(),
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
