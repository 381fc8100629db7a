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


def run_variant(rrlbm, storage, steps, warmup, tune=True, les=False, two_step=False):
    """One storage variant of the channel step on the full grid (les: RR+WALE,
    the one-launch lagged-gradient step; two_step: two steps per launch with z
    chunks of 16 planes instead of the library default, without tuning); MLUPS
    over the wall clock, the fused kernel's mean duration per step from HIP
    events on the library's stream."""
    extra = dict(steps_per_pass=2, pass_planes=16) if two_step else {}
    s = rrlbm.Solver(lx=LX, ly=LY, lz=LZ, npop=27, storage=storage, tau=TAU, force=(FORCE, 0.0, 0.0),
                     wall_y=1, wall_rr=1, use_les=int(les), cw=0.5, **extra)
    try:
        s.init_channel(U_TAU, VISC, SEED, 0.1, 0.01)
        tuning = s.autotune(trial_steps=10) if tune else None
        spp = s.params.steps_per_pass
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
    per_step = kernel_ms / 1e3 / steps
    # LES: + the velocity field (3 components in the compute type) read once and written once
    blu = 2 * 27 * STORAGE_BYTES[storage] + (6 * (8 if storage == 0 else 4) if les else 0)
    # a launch of `spp` steps moves spp * blu * n algorithmic bytes in spp * per_step
    achieved = blu * n / per_step / 1e9
    return {
        "mlups": mlups,
        "ms_per_step": wall * 1e3 / steps,
        "kernel_ms": kernel_ms / steps,
        "steps_per_launch": spp,
        "bytes_per_lu": blu,
        "achieved_gbs": achieved,
        "frac": achieved / PEAK_GBS,
        "mass_excess": d["mass_excess"],
        "ubulk_over_utau": d["u_sum"][0] / n / U_TAU,
        "tuning": tuning,
    }


def traffic_from_profiles(storage, two_step=False):
    """HBM bytes per launch from the committed rocprofv3 counter summary, if any
    (two_step: of the two-steps-per-launch kernel, whose launch covers two steps)."""
    path = os.path.join(REPO, "profiles", "fused_channel_384x128x192_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        t = json.load(fh)
    return t.get(STORAGE_NAME[storage] + ("_two_step" if two_step else ""), {}).get("hbm_bytes_per_launch")


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def oracle_mlups(make, steps, threads=None):
    """MLUPS of `steps` reference-order oracle steps on the Fields that make()
    returns, with `threads` OpenMP threads (None: the host's default)."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle

    oracle.load()
    default = oracle.get_threads()
    oracle.set_threads(threads or default)
    try:
        g, run = make(oracle)
        used = oracle.get_threads()
        t0 = time.perf_counter()
        run(g, steps)
        dt = time.perf_counter() - t0
    finally:
        oracle.set_threads(default)
    return g.n * steps / dt / 1e6, dt, used


def cpu_baseline(steps=20, steps_1t=2):
    """Oracle (C FP64, OpenMP) reference-order steps on the full grid: `steps`
    steps on all host threads (about 10 s at ~17 MLUPS on 16 threads) and
    `steps_1t` on one thread (about 10 s)."""

    def make(oracle):
        g = oracle.Fields(oracle.D3Q27_CUDA, LX, LY, LZ)
        g.init_channel_fields(U_TAU, VISC, SEED, 0.1, 0.01)
        g.init_equilibrium()
        g.set_force((FORCE, 0.0, 0.0))
        return g, lambda g, n: g.step(n, TAU, walls=(0, 1, 0), wall_rr=1)

    v, dt, used = oracle_mlups(make, steps)
    v1, dt1, _ = oracle_mlups(make, steps_1t, threads=1)
    return {"value": v, "unit": "MLUPS", "cores": used, "kind": "port",
            "sample": "%d reference-order steps (FP64 oracle, C/OpenMP) of the full %dx%dx%d grid, %.1f s on %d "
                      "threads; %d steps on 1 thread, %.1f s" % (steps, LX, LY, LZ, dt, used, steps_1t, dt1),
            "value_1_thread": v1, "cpu_model": cpu_model(), "host_threads": os.cpu_count()}


def run_ensemble(rrlbm, steps, warmup, storage=1):
    from rrlbm import configs

    inst = configs.ensemble_instances()
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


def small_ensemble(rrlbm, steps):
    """64 small periodic TG boxes (N in {32, 48}, 4.3 M cells): the regime where
    per-instance launches are launch-bound."""
    import numpy as np

    rng = np.random.default_rng(3)
    inst = []
    for _ in range(64):
        n = int(rng.choice([32, 48]))
        u0 = 0.05 / np.sqrt(3.0)
        inst.append((n, n, n, 3.0 * u0 * (n / (2 * np.pi)) / 400.0 + 0.5, u0))
    with rrlbm.Batch(inst, storage=1, keep_reference=False) as b:
        cells = b.cells()
        b.init_taylor_green_3d()
        b.step(10)
        b.step_sequential(10)
        res = {"batched": b.time(0, steps), "sequential": b.time(1, steps)}
        b.graph_capture(steps)
        b.graph_launch(1)
        res["graph"] = b.time(2, 1)
    return cells, {k: cells * steps / (v / 1e3) / 1e6 for k, v in res.items()}


def main_ensemble(args):
    import rrlbm

    rrlbm.load()
    cells, m, ke = run_ensemble(rrlbm, args.steps, args.warmup)
    scells, sm = small_ensemble(rrlbm, 200)
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
        "small_ensemble": {"workload": "64 periodic D3Q27 TG instances, N in {32, 48}, %d cells, FP32" % scells,
                           "modes_mlups": {k: round(v, 1) for k, v in sm.items()},
                           "speedup_vs_sequential": round(max(sm["graph"], sm["batched"]) / sm["sequential"], 3)},
    }
    print(json.dumps(line))


def timed(s, steps, warmup, tune=True, graph=False):
    """(MLUPS over the wall clock, mean ms per step, tuning) of `steps` fused
    steps, after autotuning (before the pack) when asked; graph=True replays a
    captured graph of `graph_len` steps instead of eager launches."""
    tuning = s.autotune(trial_steps=10) if tune else None
    s.fused_pack()
    s.fused_step(warmup)
    s.synchronize()
    if graph:
        glen = 100 if steps % 100 == 0 else steps
        s.graph_capture(glen)
        s.graph_launch(1)
        s.synchronize()
        t0 = time.perf_counter()
        ms = s.graph_launch_timed(steps // glen)
    else:
        t0 = time.perf_counter()
        ms = s.fused_step_timed(steps)
    s.synchronize()
    wall = time.perf_counter() - t0
    return s.n * steps / wall / 1e6, ms / steps, tuning


TUNING_KEYS = ("pop_group_rows", "pop_pad_bytes", "cells_per_thread", "xcd_remap", "persistent_wg_per_cu",
               "steps_per_pass", "pass_planes")


def shape_of(t):
    return None if t is None else {k: t[k] for k in TUNING_KEYS}


def line_for(metric, workload, mlups, ms, blu, dtype, steps, warmup, extra):
    ach = blu * mlups * 1e6 / 1e9
    d = {"metric": metric, "value": round(mlups, 1), "unit": "MLUPS", "n_gpus": 1, "steps": steps, "warmup": warmup,
         "ms_per_step": round(ms, 5), "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": dtype,
         "data": "synthetic", "config": {"workload": workload},
         "roofline": {"bound": "hbm", "achieved": round(ach, 1), "peak": PEAK_GBS, "unit": "GB/s",
                      "frac": round(ach / PEAK_GBS, 4), "traffic": None, "bytes_per_lu": blu}}
    d.update(extra)
    return d


def tg_cpu_baseline(steps=2000):
    """The FP64 oracle (C/OpenMP) over the full C1 run: 2 000 steps of 100x100x4."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import math

    import oracle

    u0 = 0.01 / math.sqrt(3.0)
    nu = u0 * 100 / 100.0
    g = oracle.Fields(oracle.D3Q27_CUDA, 100, 100, 4)
    g.init_tg_fields(u0)
    g.init_equilibrium()
    g.set_force((0.0, 0.0, 0.0))
    t0 = time.perf_counter()
    g.step(steps, 3 * nu + 0.5, walls=(0, 0, 0), wall_rr=0)
    dt = time.perf_counter() - t0
    return {"value": round(40000 * steps / dt / 1e6, 2), "unit": "MLUPS", "cores": oracle.get_threads(),
            "kind": "port", "sample": "%d steps of 100x100x4 (FP64 oracle, C/OpenMP), %.1f s" % (steps, dt),
            "cpu_model": cpu_model()}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def main_configs(args):
    """C4 LES channel; C3 grid with LES, BGK, sixth-order RR and FDBC; C2 cavity; C1 Taylor-Green:
    one JSON line each."""
    import rrlbm
    from rrlbm import configs

    rrlbm.load()
    out = []

    def les_c4(oracle):
        nu = 0.0045 * 32 / 180.0
        g = oracle.Fields(oracle.D3Q27_CUDA, 192, 64, 96)
        g.init_channel_fields(0.0045, nu, SEED, 0.1, 0.01)
        g.init_equilibrium()
        g.set_force((0.0045 ** 2 / 32, 0.0, 0.0))
        return g, lambda g, n: g.step(n, 3 * nu + 0.5, use_les=True, walls=(0, 1, 0), wall_rr=1)

    def cavity(q):
        def make(oracle):
            kind = oracle.D3Q27_CUDA if q == 27 else oracle.D3Q19
            g = oracle.Fields(kind, 128, 128, 128)
            g.init_equilibrium()
            tau = 3 * 0.1 * 128 / 1000.0 + 0.5
            return g, lambda g, n: g.step(n, tau, walls=(1, 1, 1), wall_rr=1, lid_face=3, lid_u=(0.1, 0.0, 0.0))
        return make

    def cpu_line(make, steps, what):
        v, dt, used = oracle_mlups(make, steps)
        return {"value": round(v, 2), "unit": "MLUPS", "cores": used, "kind": "port",
                "sample": "%d reference-order steps of %s (FP64 oracle, C/OpenMP), %.1f s" % (steps, what, dt),
                "cpu_model": cpu_model()}

    cpu_c4 = cpu_line(les_c4, 10, "the 192x64x96 LES channel")
    cpu_c2 = {q: cpu_line(cavity(q), 5, "the 128^3 D3Q%d cavity" % q) for q in (19, 27)}
    # C4: channel LES-WALE 192x64x96, lagged gradient (one launch per step) and reference order
    # (1.2 M cells: each step is ~30-50 us, so the steps run as graph replays)
    for st in (2, 1, 0):
        for lagged in (1, 0):
            s, meta = configs.channel_les(storage=st, les_lagged=lagged)
            mlups, ms, tun = timed(s, args.steps, args.warmup, tune=not args.no_tune, graph=not args.no_graph)
            s.close()
            log("C4 storage %d lagged %d: %.1f MLUPS" % (st, lagged, mlups))
            blu = 2 * 27 * STORAGE_BYTES[st] + 6 * (8 if st == 0 else 4) + (0 if lagged else 27 * STORAGE_BYTES[st])
            out.append(line_for("MLUPS, channel LES-WALE Re_tau=180 192x64x96 D3Q27", "C4 LES %s %s" % (
                STORAGE_NAME[st], "lagged" if lagged else "reference-order"), mlups, ms, blu, STORAGE_NAME[st],
                args.steps, args.warmup, {"tuning": shape_of(tun), "launch": "eager" if args.no_graph else "graph",
                                          "cpu_baseline": cpu_c4}))
    # C3 grid with WALE (the DNS grid, LES on): the fused RR+WALE kernel at full size
    for st in (2, 1):
        s, meta = configs.channel(storage=st, les=True)
        mlups, ms, tun = timed(s, args.steps, args.warmup, tune=not args.no_tune)
        s.close()
        log("C3 grid LES storage %d: %.1f MLUPS" % (st, mlups))
        out.append(line_for("MLUPS, channel RR+WALE Re_tau=180 384x128x192 D3Q27", "C3 grid LES %s lagged"
                            % STORAGE_NAME[st], mlups, ms, 2 * 27 * STORAGE_BYTES[st] + 24, STORAGE_NAME[st],
                            args.steps, args.warmup, {"tuning": shape_of(tun)}))
    # C3 grid under the BGK baseline operator (A/B against the RR step on the same grid)
    for st in (1, 2):
        s, meta = configs.channel_bgk(storage=st)
        mlups, ms, tun = timed(s, args.steps, args.warmup, tune=not args.no_tune)
        s.close()
        log("C3 grid BGK storage %d: %.1f MLUPS" % (st, mlups))
        out.append(line_for("MLUPS, channel BGK Re_tau=180 384x128x192 D3Q27", "C3 grid BGK %s" % STORAGE_NAME[st],
                            mlups, ms, 2 * 27 * STORAGE_BYTES[st], STORAGE_NAME[st], args.steps, args.warmup,
                            {"tuning": shape_of(tun)}))
    # C3 grid with the fourth- to sixth-order RR terms
    for st in (1, 2):
        s, meta = configs.channel(storage=st, rr_order=6)
        mlups, ms, tun = timed(s, args.steps, args.warmup, tune=not args.no_tune)
        s.close()
        log("C3 grid RR order 6 storage %d: %.1f MLUPS" % (st, mlups))
        out.append(line_for("MLUPS, channel RR (orders 4-6 on) Re_tau=180 384x128x192 D3Q27",
                            "C3 grid RR6 %s" % STORAGE_NAME[st], mlups, ms, 2 * 27 * STORAGE_BYTES[st],
                            STORAGE_NAME[st], args.steps, args.warmup, {"tuning": shape_of(tun)}))
    # C3 grid with the FDBC wall reconstruction in place of RRBC (a row pass + the step)
    for st in (1, 2):
        s, meta = configs.channel(storage=st, wall_rr=2)
        mlups, ms, tun = timed(s, args.steps, args.warmup, tune=not args.no_tune)
        s.close()
        log("C3 grid FDBC storage %d: %.1f MLUPS" % (st, mlups))
        out.append(line_for("MLUPS, channel RR + FDBC walls Re_tau=180 384x128x192 D3Q27",
                            "C3 grid FDBC %s" % STORAGE_NAME[st], mlups, ms, 2 * 27 * STORAGE_BYTES[st],
                            STORAGE_NAME[st], args.steps, args.warmup, {"tuning": shape_of(tun)}))
    # C2: cavity 128^3 Re 1000, D3Q19 and D3Q27
    for q in (19, 27):
        for st in (1, 0):
            s, meta = configs.cavity(storage=st, npop=q)
            mlups, ms, tun = timed(s, args.steps, args.warmup, tune=not args.no_tune)
            s.close()
            log("C2 D3Q%d storage %d: %.1f MLUPS" % (q, st, mlups))
            out.append(line_for("MLUPS, lid-driven cavity 128^3 Re=1000", "C2 cavity D3Q%d %s" % (q, STORAGE_NAME[st]),
                                mlups, ms, 2 * q * STORAGE_BYTES[st], STORAGE_NAME[st], args.steps, args.warmup,
                                {"tuning": shape_of(tun), "cpu_baseline": cpu_c2[q]}))
    # C1: Taylor-Green 100x100x4 (40 000 cells: launch-bound), full 2 000 steps;
    # the CPU oracle runs the same 2 000 reference-order steps
    cpu = tg_cpu_baseline()
    for st in (0, 1):
        s, meta = configs.taylor_green_2d(storage=st)
        mlups, ms, _ = timed(s, 2000, 20, tune=not args.no_tune)
        # 40 000 cells are launch-bound: the same 2 000 steps as 20 replays of a
        # captured graph of 100 steps
        if args.no_graph:
            gms, gmlups = ms * 2000, mlups
            s.close()
        else:
            s.graph_capture(100)
            s.graph_launch(1)
            s.synchronize()
            t0 = time.perf_counter()
            gms = s.graph_launch_timed(20)
            gwall = time.perf_counter() - t0
            s.close()
            gmlups = s.n * 2000 / gwall / 1e6
        log("C1 storage %d: %.1f MLUPS eager, %.1f MLUPS graph" % (st, mlups, gmlups))
        out.append(line_for("MLUPS, Taylor-Green 100x100x4", "C1 TG %s" % STORAGE_NAME[st], max(mlups, gmlups),
                            min(ms, gms / 2000), 2 * 27 * STORAGE_BYTES[st], STORAGE_NAME[st], 2000, 20,
                            {"cpu_baseline": cpu, "eager_mlups": round(mlups, 1), "graph_mlups": round(gmlups, 1)}))
    for d in out:
        print(json.dumps(d))


def main_les_step(args):
    """The C3-grid (384 x 128 x 192) fused steps the LES analysis compares, with
    the library defaults and eager launches: FP16 RR+WALE (lagged), FP32
    RR+WALE (lagged) and the FP16 RR step without LES; one JSON line each."""
    import rrlbm
    from rrlbm import configs

    rrlbm.load()
    for st, les in ((2, True), (1, True), (2, False)):
        s, meta = configs.channel(storage=st, les=les)
        mlups, ms, _ = timed(s, args.steps, args.warmup, tune=False)
        s.close()
        blu = 2 * 27 * STORAGE_BYTES[st] + (24 if les else 0)
        print(json.dumps(line_for("MLUPS, channel %s Re_tau=180 384x128x192 D3Q27" % ("RR+WALE" if les else "RR"),
                                  "C3 grid %s %s" % ("LES lagged" if les else "DNS", STORAGE_NAME[st]), mlups, ms,
                                  blu, STORAGE_NAME[st], args.steps, args.warmup, {})), flush=True)


def main_validate(args):
    """Precision checks over long runs: FP32 / FP16 storage against the FP64
    run of the same kernels (and a second-seed FP64 run as the chaotic floor)."""
    import numpy as np

    import rrlbm
    from rrlbm import configs

    rrlbm.load()
    res = {}
    # C1: Taylor-Green energy decay after 2000 steps
    ke = {}
    for st in (0, 1, 2):
        s, meta = configs.taylor_green_2d(storage=st)
        k0 = s.diag_flow()["kinetic_energy"]
        s.fused_pack()
        s.fused_step(1999)
        ke[st] = s.fused_diag()["kinetic_energy"] / k0
        s.close()
    log("C1 done")
    analytic = float(np.exp(-4 * meta["k"] ** 2 * meta["nu"] * 2000))
    orc = np.loadtxt(os.path.join(REPO, "tests", "golden", "tg2d_c1_ke_oracle.txt"))
    orc_ratio = float(orc[-1, 1] / orc[0, 1])
    res["C1_taylor_green"] = {"ke_ratio_f64": ke[0], "ke_ratio_oracle": orc_ratio, "analytic": analytic,
                              "rel_err_f64_vs_oracle": abs(ke[0] / orc_ratio - 1),
                              "rel_err_f64_vs_analytic": abs(ke[0] / analytic - 1),
                              "rel_err_f32_vs_f64": abs(ke[1] / ke[0] - 1), "rel_err_f16_vs_f64": abs(ke[2] / ke[0] - 1)}
    # C2: cavity 128^3 centre lines after 20 000 steps, FP32 vs FP64
    for q in (19, 27):
        lines = {}
        for st in (0, 1):
            s, meta = configs.cavity(storage=st, npop=q)
            s.fused_pack()
            s.fused_step(args.cavity_steps - 1)
            s.fused_unpack()
            n = s.lx
            U = s.download(rrlbm.Field.UX).reshape(n, n, n)
            V = s.download(rrlbm.Field.UY).reshape(n, n, n)
            c = n // 2
            lines[st] = (0.25 * (U[c - 1, :, c - 1] + U[c - 1, :, c] + U[c, :, c - 1] + U[c, :, c]),
                         0.25 * (V[c - 1, c - 1, :] + V[c - 1, c, :] + V[c, c - 1, :] + V[c, c, :]))
            s.close()
            log("C2 D3Q%d storage %d done" % (q, st))
        res["C2_cavity_D3Q%d" % q] = {"steps": args.cavity_steps,
                                      "max_abs_du_over_lid": float(np.abs(lines[1][0] - lines[0][0]).max() / 0.1),
                                      "max_abs_dv_over_lid": float(np.abs(lines[1][1] - lines[0][1]).max() / 0.1),
                                      "u_min_f64": float(lines[0][0].min())}

    # C3 / C4: channel mean U+ profiles
    def channel_stats(make, st, seed, steps, every, window):
        s, meta = make(storage=st, seed=seed)
        s.fused_pack()
        acc, cnt, done = None, 0, 1
        while done < steps:
            s.fused_step(min(every, steps - done))
            done = min(done + every, steps)
            if done > steps - window:
                s.fused_unpack()
                p = s.output_profile()
                acc = p.copy() if acc is None else acc + p
                cnt += 1
        d = s.fused_diag()
        s.close()
        log("channel %dx%dx%d storage %d seed %d: %d steps" % (s.lx, s.ly, s.lz, st, seed, steps))
        prof = acc / cnt
        return prof[:, 0] / meta["u_tau"], np.sqrt(np.maximum(prof[:, 3], 0)) / meta["u_tau"], d["u_sum"][0] / d[
            "ncells"] / meta["u_tau"]

    for name, make, steps, window in (("C4_channel_les", configs.channel_les, args.les_steps, 20000),
                                      ("C3_channel_dns", configs.channel, args.dns_steps, 10000)):
        base = channel_stats(make, 0, 2024, steps, 1000, window)
        other = channel_stats(make, 0, 7, steps, 1000, window)
        r = {"steps": steps, "window": window, "ubulk_plus_f64": base[2],
             "umean_plus_f64_seedB_rel": float(np.abs(other[0] - base[0]).max() / base[0].max())}
        for st in (1, 2):
            o = channel_stats(make, st, 2024, steps, 1000, window)
            r["umean_plus_%s_rel" % STORAGE_NAME[st]] = float(np.abs(o[0] - base[0]).max() / base[0].max())
            r["urms_plus_%s_abs" % STORAGE_NAME[st]] = float(np.abs(o[1] - base[1]).max())
            r["ubulk_plus_%s" % STORAGE_NAME[st]] = o[2]
        res[name] = r
    print(json.dumps({"workload": "validate", "results": res}))


def glibc_rand():
    """(srand, draw) of the C library's generator: the reference draws its
    forcing_GP phases as rand() / RAND_MAX after srand(12345) (main.cu:3-5,30)."""
    import ctypes

    libc = ctypes.CDLL(None)
    libc.srand.argtypes = [ctypes.c_uint]
    libc.rand.restype = ctypes.c_int
    rand_max = 2147483647  # glibc RAND_MAX
    return libc.srand, (lambda: libc.rand() / rand_max)


def main_turbulence(args):
    """Long run of a turbulent channel (fused path) with the reference's
    protocol (main.cu:29-84): init_arrays (initial.cu:268-325), forcing_GP with
    a new phase rand()/RAND_MAX every 2000 steps below step 50 000 (srand of
    --turb-phase-seed, 12345 in the reference), then the uniform force.  Every
    1000 steps (NDIAG, parameters.h:17) the bulk velocity U_b+ of the diag_flow
    sums; U_b+ as the reference's notebook averages it (the last 3000 diag
    samples, turbulent.ipynb:78-106) and over the final window; the mean U+ and
    the rms profiles (u', v', w', -u'v' in wall units, the two halves folded)
    over the window, against the spectral Re_tau = 180 profile (Re180.prof).
    The grid defaults to 384 x 128 x 192 at u_tau = 0.0045; --turb-visc sets
    the viscosity and u_tau = 2 Re_tau visc / LY as globals.cu:57-63 does (the
    reference's 64^3 run: --turb-grid 64x64x64 --turb-visc 0.000622)."""
    import numpy as np

    import rrlbm
    from rrlbm import configs

    rrlbm.load()
    spec = np.loadtxt(os.path.join(REPO, "tests", "golden", "re180_spectral_uplus.txt"))
    lx, ly, lz = (int(v) for v in args.turb_grid.split("x"))
    re_tau = 180.0
    u_tau = 2.0 * re_tau * args.turb_visc / ly if args.turb_visc > 0 else U_TAU
    extra = {}
    if args.turb_les:
        extra["les_lagged"] = 0 if args.turb_les_exact else 1
    s, meta = configs.channel(storage=args.turb_storage, lx=lx, ly=ly, lz=lz, u_tau=u_tau, les=bool(args.turb_les),
                              **extra)
    nu = meta["nu"]
    s.init_arrays(u_tau, nu / u_tau)
    srand, draw = glibc_rand()
    srand(args.turb_phase_seed)
    gp_end, phase = args.turb_gp_steps, 0.0
    ndiag = 1000
    t0 = time.perf_counter()
    ub = []  # (step, U_b+) every ndiag steps

    def diag(step):
        d = s.fused_diag()
        ub.append((step, d["u_sum"][0] / d["ncells"] / u_tau))

    # step 1: forcing_GP(1) then its collision (pack); steps 2 .. gp_end-1 in
    # the kernel's forcing_GP mode; then the uniform force
    s.forcing_GP(1, phase)
    s.fused_pack()
    istep = 1
    while istep + 1 < min(gp_end, args.turb_steps + 1):
        istep += 1
        if istep % 2000 == 0:
            phase = draw()
        s.fused_step_gp(istep, phase)
        if istep % ndiag == 0:
            diag(istep)
        if istep % 10000 == 0:
            log("turbulence: forcing_GP step %d, %.0f s" % (istep, time.perf_counter() - t0))
    acc, cnt = None, 0
    while istep < args.turb_steps:
        n = min(ndiag - istep % ndiag, args.turb_steps - istep)
        s.fused_step(n)
        istep += n
        if istep % ndiag == 0:
            diag(istep)
        if istep % 50000 == 0:
            log("turbulence: %d steps, %.0f s" % (istep, time.perf_counter() - t0))
        if istep > args.turb_steps - args.turb_window and istep % ndiag == 0:
            s.fused_unpack()
            p = s.output_profile()
            acc = p.copy() if acc is None else acc + p
            cnt += 1
    s.close()
    ubv = np.array([u for _, u in ub])
    prof = acc / cnt

    def fold(col, sign=1.0):
        return 0.5 * (prof[:, col] + sign * prof[::-1, col])[: ly // 2]

    u = fold(0)
    yw = np.arange(ly // 2) + 0.5
    yplus = yw * u_tau / nu
    uplus = u / u_tau
    ref = np.interp(yplus, spec[:, 1], spec[:, 2])
    band = (yplus > 5) & (yplus < 150)
    dudy = (9 * u[0] - u[1]) / 3.0      # quadratic through u = 0 at the half-way wall
    rnd = lambda a: [round(float(v), 5) for v in a]
    res = {"workload": "turbulence", "grid": [lx, ly, lz], "storage": STORAGE_NAME[args.turb_storage],
           "les": ("exact" if args.turb_les_exact else "lagged") if args.turb_les else None,
           "u_tau": u_tau, "nu": nu, "steps": args.turb_steps, "gp_steps": gp_end, "phase_seed": args.turb_phase_seed,
           "window": args.turb_window, "samples": cnt, "seconds": round(time.perf_counter() - t0, 1),
           "ubulk_plus_last3000_diag": float(ubv[-3000:].mean()), "ubulk_plus_window": float(prof[:, 0].mean() / u_tau),
           "ubulk_plus_spectral": 15.709, "u_tau_from_wall_gradient": float(np.sqrt(nu * dudy)),
           "max_rel_dUplus_5_150_vs_spectral": float((np.abs(uplus - ref) / ref)[band].max()),
           "centre_uplus": float(uplus[-1]), "centre_uplus_spectral": float(spec[-1, 2]),
           "profile": {"yplus": rnd(yplus), "uplus": rnd(uplus),
                       "urms": rnd(np.sqrt(np.maximum(fold(3), 0)) / u_tau),
                       "vrms": rnd(np.sqrt(np.maximum(fold(4), 0)) / u_tau),
                       "wrms": rnd(np.sqrt(np.maximum(fold(5), 0)) / u_tau),
                       "uv": rnd(-fold(6, -1.0) / u_tau ** 2)},
           "ubulk_plus_series": [[int(k), round(float(v), 5)] for k, v in ub]}
    line = json.dumps(res)
    if args.turb_out:
        with open(args.turb_out, "w") as fh:
            fh.write(line + "\n")
    summary = {k: v for k, v in res.items() if k not in ("profile", "ubulk_plus_series")}
    print(json.dumps(summary))


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


def main_sizes(args):
    """The fused step's cost as a function of grid size (library defaults, eager
    launches, best of two contexts): FP16 RR+WALE, FP16 RR and FP32 RR+WALE on
    channel grids of 0.98 M to 9.4 M cells, and a least-squares fit
    us_per_step = t0 + c * cells (t0: the fixed cost of a launch on one stream;
    c: the sustained cost per cell, whose bytes per lattice update / c is the
    rate the kernel reaches on large grids)."""
    import numpy as np

    import rrlbm
    from rrlbm import configs

    rrlbm.load()
    grids = [(192, 64, 80), (192, 64, 96), (192, 64, 128), (384, 64, 96), (384, 128, 96), (384, 128, 192)]
    for st, les in ((2, True), (2, False), (1, True), (1, False)):
        blu = 2 * 27 * STORAGE_BYTES[st] + (24 if les else 0)
        pts = []
        for lx, ly, lz in grids:
            n = lx * ly * lz
            best = 1e9
            for _ in range(2):
                s, _meta = configs.channel(storage=st, lx=lx, ly=ly, lz=lz, les=les)
                s.fused_pack()
                s.fused_step(args.warmup)
                s.synchronize()
                best = min(best, s.fused_step_timed(args.steps) / args.steps)
                s.close()
            us = best * 1e3
            pts.append((n, us))
            print(json.dumps({"storage": STORAGE_NAME[st], "les": les, "grid": [lx, ly, lz], "cells": n,
                              "us_per_step": round(us, 2), "frac": round(n * blu / (us * 1e-6) / (PEAK_GBS * 1e9), 4)}),
                  flush=True)
        n_arr = np.array([p[0] for p in pts], dtype=float)
        t_arr = np.array([p[1] for p in pts])
        c, t0 = np.polyfit(n_arr, t_arr, 1)
        print(json.dumps({"storage": STORAGE_NAME[st], "les": les, "fit_t0_us": round(float(t0), 2),
                          "fit_ns_per_kcell": round(float(c) * 1e6, 3),
                          "sustained_frac": round(blu / (c * 1e-6) / (PEAK_GBS * 1e9), 4)}), flush=True)


def main_ceiling():
    """What plain streams reach on this GPU, for comparison with the fused step:
    torch's device-to-device copy (hipMemcpyAsync), a read-only reduction and a
    write-only fill, each over 1 GiB buffers, median of 20 timed repetitions
    (HIP events on torch's stream).  Bytes counted: read + written."""
    import torch

    n = (1 << 30) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    out = {}
    for name, fn, moved in (("copy", lambda: b.copy_(a), 8 * n), ("read_sum", lambda: torch.sum(a), 4 * n),
                            ("write_fill", lambda: b.fill_(1.0), 4 * n),
                            ("elementwise_add", lambda: torch.add(a, 1.0, out=b), 8 * n)):
        fn()
        fn()
        torch.cuda.synchronize()
        ts = []
        for _ in range(20):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            fn()
            e1.record()
            e1.synchronize()
            ts.append(e0.elapsed_time(e1))
        ms = sorted(ts)[len(ts) // 2]
        out[name] = {"us": round(ms * 1e3, 1), "tb_per_s": round(moved / (ms * 1e-3) / 1e12, 3)}
    print(json.dumps({"workload": "ceiling", "buffer_bytes": 4 * n, "results": out}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-les", action="store_true", help="channel workload: leave out the RR+WALE variants")
    ap.add_argument("--no-tune", action="store_true", help="library default layout and launch shape (no autotuning)")
    ap.add_argument("--two-step", action="store_true",
                    help="channel workload: FP32/FP64 variants with two steps per launch (z chunks of 16 planes), "
                         "no autotuning of those variants")
    ap.add_argument("--no-graph", action="store_true", help="configs workload: eager launches only (no HIP graphs)")
    ap.add_argument("--variants", default="1,2,0", help="storage variants; the first is the headline")
    ap.add_argument("--workload", default="channel",
                    choices=["channel", "ensemble", "configs", "validate", "copy", "ceiling", "sizes", "turbulence", "les-step"])
    ap.add_argument("--turb-steps", type=int, default=1000000)
    ap.add_argument("--turb-window", type=int, default=300000)
    ap.add_argument("--turb-gp-steps", type=int, default=50000)
    ap.add_argument("--turb-storage", type=int, default=1)
    ap.add_argument("--turb-grid", default="%dx%dx%d" % (LX, LY, LZ))
    ap.add_argument("--turb-visc", type=float, default=0.0,
                    help="viscosity; u_tau = 2 Re_tau visc / LY (globals.cu:57-63); 0: u_tau = 0.0045")
    ap.add_argument("--turb-les", type=int, default=0)
    ap.add_argument("--turb-les-exact", action="store_true", help="LES in the reference order (velocity pass first)")
    ap.add_argument("--turb-phase-seed", type=int, default=12345)
    ap.add_argument("--turb-out", default="", help="write the full result (profiles, U_b+ series) to this file")
    ap.add_argument("--cavity-steps", type=int, default=20000)
    ap.add_argument("--les-steps", type=int, default=50000)
    ap.add_argument("--dns-steps", type=int, default=30000)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the solver has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "ensemble":
        return main_ensemble(args)
    if args.workload == "copy":
        return main_copy()
    if args.workload == "ceiling":
        return main_ceiling()
    if args.workload == "sizes":
        return main_sizes(args)
    if args.workload == "configs":
        return main_configs(args)
    if args.workload == "turbulence":
        return main_turbulence(args)
    if args.workload == "validate":
        return main_validate(args)
    if args.workload == "les-step":
        return main_les_step(args)
    import rrlbm

    rrlbm.load()
    variants = [int(v) for v in args.variants.split(",")]
    res = {st: run_variant(rrlbm, st, args.steps, args.warmup, tune=not args.no_tune,
                           two_step=args.two_step and st != 2) for st in variants}
    les = {} if args.no_les else {st: run_variant(rrlbm, st, args.steps, args.warmup, tune=not args.no_tune, les=True)
                                  for st in (1, 2)}
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
                     "frac": round(head["frac"], 4),
                     "traffic": traffic_from_profiles(st0, head["steps_per_launch"] == 2),
                     "bytes_per_lu": head["bytes_per_lu"], "kernel_ms": round(head["kernel_ms"], 5),
                     "steps_per_launch": head["steps_per_launch"],
                     "launch_ms": round(head["kernel_ms"] * head["steps_per_launch"], 5)},
        "variants": {STORAGE_NAME[st]: {"mlups": round(r["mlups"], 1), "frac": round(r["frac"], 4),
                                         "achieved_gbs": round(r["achieved_gbs"], 1),
                                         "bytes_per_lu": r["bytes_per_lu"], "kernel_ms": round(r["kernel_ms"], 5),
                                         "ubulk_over_utau": round(r["ubulk_over_utau"], 4),
                                         "steps_per_launch": r["steps_per_launch"],
                                         "tuning": shape_of(r["tuning"])}
                     for st, r in res.items()},
        # the same grid with WALE LES (lagged gradient, one launch per step)
        "variants_les": {STORAGE_NAME[st]: {"mlups": round(r["mlups"], 1), "frac": round(r["frac"], 4),
                                             "achieved_gbs": round(r["achieved_gbs"], 1),
                                             "bytes_per_lu": r["bytes_per_lu"], "kernel_ms": round(r["kernel_ms"], 5),
                                             "tuning": shape_of(r["tuning"])}
                         for st, r in les.items()},
        "cpu_baseline": None if args.no_cpu else cpu_baseline(),
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
