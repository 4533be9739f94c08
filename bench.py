#!/usr/bin/env python3
"""Detected coincidences/s of the MI355X photon-transport library on the head workload.

Workload (one step = one full simulation): 128x128x96 head phantom at 2 mm (8 tissues),
18F activity grey 4 : white 1 : other 0.5, 1e9 expected decays over a 10 s acquisition,
cylindrical scanner R = 40 cm, H = 25 cm, 12% energy resolution, 350-600 keV window,
3D sinogram 168 x 168 x 1293 (int32), emission images, energy spectrum. Inputs are
resident on the device before timing; every step clears the tallies and simulates all
decays with a new seed.

Prints one JSON line: metric, value, unit, timing, config, a roofline object for the dominant
kernel and (head and torso workloads) the CPU baseline. The head line also carries the second
half of the headline metric, `time_to_5pct_nrmse`: the time Monte Carlo and the learned
surrogate take to a 5 %-NRMSE scatter sinogram on one phantom of each evaluation family, with
the U-Net's MFMA roofline (--no-surrogate leaves it out).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "mcgpu-pet_amd"))

MATERIALS = os.path.join(ROOT, "tests", "golden", "sample_simulation", "materials")

# Algorithmic work units (VALU lane-ops, HBM bytes) per event, counted from the reference
# source: RANECU draw 25 ops; Woodcock step 83 ops incl. its 2 draws, 4 B packed voxel word;
# material/energy change 24 B of (a,b); Compton 40 + 35 per shell evaluation; detected
# coincidence 32 B of atomic read-modify-write; with the voxel dose tally, 32 B (two u64
# read-modify-writes of the voxel's 16-byte record) per Compton or photoelectric deposit.
U_DRAW, U_STEP, U_COMPTON, U_SHELL = 25, 83, 40, 35
B_STEP, B_MATCHANGE, B_COINC, B_DOSE = 4, 24, 32, 32
VALU_PEAK = 39.3e12      # 256 CU x 4 SIMD x 16 lanes x 2.4 GHz, non-packed lane-ops/s
HBM_PEAK = 8.0e12        # B/s


def algorithmic(stats, voxel_dose=False):
    steps = stats["woodcock_steps"]
    other_draws = max(0, stats["rn_draws"] - 2 * steps)
    ops = (U_STEP * steps + U_DRAW * other_draws + U_COMPTON * stats["compton"] + U_SHELL * stats["compton_shell_evals"])
    bytes_ = B_STEP * steps + B_MATCHANGE * stats["material_changes"] + B_COINC * stats["coincidences"]
    if voxel_dose:
        bytes_ += B_DOSE * (stats["compton"] + stats["photoelectric"])
    return ops, bytes_


def load_traffic(name="transport_head_traffic.json", key="hbm_bytes_per_launch"):
    """Measured HBM bytes (calibrated FETCH_SIZE + WRITE_SIZE) from a committed profile."""
    p = os.path.join(ROOT, "profiles", name)
    if os.path.exists(p):
        with open(p) as f:
            return json.load(f).get(key)
    return None


def host_description():
    """The CPU baseline's host: CPUs this process may run on, the CPU model, the C compiler."""
    import subprocess
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        gcc = subprocess.run(["gcc", "--version"], capture_output=True, text=True).stdout.splitlines()[0]
    except (OSError, IndexError):
        gcc = None
    return {"affinity_cpus": len(os.sched_getaffinity(0)), "os_cpu_count": os.cpu_count(), "cpu_model": model,
            "compiler": gcc, "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


def baseline_threads(requested):
    """Threads of the CPU baseline: the CPUs this process may run on, capped by the host's
    per-job CPU share when the environment states one (OMP_NUM_THREADS)."""
    if requested > 0:
        return requested
    n = len(os.sched_getaffinity(0))
    share = os.environ.get("OMP_NUM_THREADS", "")
    if share.isdigit() and int(share) > 0:
        n = min(n, int(share))
    return max(1, n)


def cpu_baseline(cfg, ph, vs, names, threads, stride, e_max=0.0):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py
    from mcpet import phantoms
    tab = oracle_py.OracleTables(names + [""] * (15 - len(names)), ph.density_max(), "libm", e_max=e_max)
    o = oracle_py.Oracle(phantoms.config_dict(cfg), ph.voxel_mat_dens(), ph.shape, vs, tab, "libm")
    r = o.run(nthreads=threads, voxel_stride=stride)
    c = r["counters"]
    return {"value": c["coincidences"] / r["wall_seconds"], "unit": "coincidences/s", "cores": threads, "kind": "port",
            "sample": f"every {stride}th voxel of the same phantom, full 10 s acquisition: "
                      f"{c['decays']} decays, "
                      f"{c['coincidences']} coincidences in {r['wall_seconds']:.1f} s (C restatement of the reference "
                      f"transport built with gcc -O3 -march=x86-64-v3, system libm, {threads} threads)",
            "host": host_description()}


def singles_roofline(steps, acq_ps):
    """HBM roofline of the singles pipeline (time sort + pairing, coinc.hip) from its HIP-event
    time. Algorithmic bytes per single: the key/index build reads the 32-byte record and writes
    an 8-byte key and a 4-byte index (44 B); every radix pass reads the keys for the digit
    histogram (8 B) and reads and writes key and index in the scatter (2 x 12 B), 32 B a pass;
    the pairing reads a single's key and index and its neighbours' keys (about 20 B)."""
    bits = int(acq_ps + (1 << 40)).bit_length()
    passes = (bits + 7) // 8
    per_single = 44 + 32 * passes + 20
    n = sum(x for x, _, _ in steps) / len(steps)
    ms = sum(t for _, t, _ in steps) / len(steps)
    achieved = n * per_single / (ms * 1e-3)
    classes = {k: sum(c[k] for _, _, c in steps) / len(steps) for k in steps[0][2]}
    return {"singles_per_step": n, "coincidences_per_step": classes, "sort_and_pairing_ms": ms,
            "singles_per_s": n / (ms * 1e-3), "radix_passes": passes,
            "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                         "frac": achieved / HBM_PEAK, "algorithmic_bytes_per_single": per_single,
                         "kernels": "k_keys_from_singles, k_rs_hist, k_scan_tile, k_scan_add, k_rs_scatter, k_pair"}}


def bench_batch(a):
    """Many-phantom workload: a.batch_size varied phantoms (random_phantoms: 32-160 x 32-160 x
    32-96 voxels at 2 mm, 1e6-1e8 photon histories each, 10 s), sinogram 168x168x1293 each.
    Rates of three ways to run it: the plain kernel launched per phantom, the persistent
    kernel launched per phantom, and the batched persistent kernel (one launch for all
    phantoms, replayed as a HIP graph with the autotuned launch shape and phase policy)."""
    import mcpet
    from mcpet import phantoms
    specs = phantoms.random_phantoms(a.batch_size)
    names = phantoms.head_materials(MATERIALS)
    cfg = phantoms.make_config(seed=20261202, threads_per_block=a.tpb, acquisition_time_s=10.0, nrad=168, materials=names)
    phs = [mcpet.phantom_from_arrays(*sp) for sp in specs]
    dmax = np.max([p.density_max() for p in phs], axis=0)
    tab = mcpet.load_materials(names, dmax)

    def per_phantom(kernel):
        coinc, secs = 0, 0.0
        for ph in phs:
            with mcpet.Simulation(cfg, ph, tab) as sim:
                sim.run(kernel=kernel, seed=1000)      # warm-up of this context
                sim.sync()
                t0 = time.perf_counter()
                sim.run(kernel=kernel, seed=2000)
                st = sim.stats()
                secs += time.perf_counter() - t0
                coinc += st["coincidences"]
        return coinc, secs

    plain_c, plain_s = per_phantom(mcpet.KERNEL_PLAIN)
    pers_c, pers_s = per_phantom(mcpet.KERNEL_PERSISTENT)
    with mcpet.Batch(cfg, phs, tab) as b:
        tune = b.autotune(0.1)
        for w in range(a.warmup):
            b.run(seed=1000 + w, use_graph=True)
            b.sync()
        coinc, decays, kms = 0, 0, []
        t0 = time.perf_counter()
        for s in range(a.steps):
            b.run(seed=2000 + s, use_graph=True)
            b.sync()
            sts = [b.stats(k) for k in range(len(phs))]
            coinc += sum(x["coincidences"] for x in sts)
            decays += sum(x["decays"] for x in sts)
            kms.append(sts[0]["kernel_ms"])
        elapsed = time.perf_counter() - t0
    value = coinc / elapsed
    out = {"metric": "detected coincidences/s on 1 MI355X", "value": value, "unit": "coincidences/s", "n_gpus": 1,
           "steps": a.steps, "warmup": a.warmup, "ms_per_step": 1e3 * elapsed / a.steps, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": f"{len(phs)} varied phantoms (32-160 x 32-160 x 32-96 voxels at 2 mm, 1e6-1e8 "
                                  "photon histories each, 10 s), sinogram 168x168x1293 per phantom, one batched "
                                  "persistent launch replayed as a HIP graph", "decays_per_step": decays / a.steps,
                      "threads_per_block": a.tpb, "parallelism": "single GPU"},
           "kernel_ms_avg": sum(kms) / len(kms),
           "plain_per_phantom": {"value": plain_c / plain_s, "seconds": plain_s},
           "persistent_per_phantom": {"value": pers_c / pers_s, "seconds": pers_s},
           "speedup_vs_plain_per_phantom": value / (plain_c / plain_s),
           "autotune": tune}
    print(json.dumps(out))


EVAL_FAMILIES = {
    "scatter": "scatter_phantoms 128x128x64 at 2 mm, the training family",
    "head": "head phantoms 128x128x96 at 2 mm, seed-jittered, not seen in training",
    "batch": "random_phantoms 32-160 x 32-160 x 32-96 at 2 mm, not seen in training",
}


def eval_phantoms(family, n, seed):
    """n phantoms of a family, activity scaled to 1e9 expected decays per second."""
    from mcpet import surrogate as S
    return S.family_phantoms(family, n, seed)


def time_to_nrmse(specs, est, steps, warmup, max_ref_runs=8, log=None):
    """Time to a 5 %-NRMSE scatter sinogram on the 32 x 56 x 56 grid, Monte Carlo against the
    surrogate (low-count run + device rebinning + U-Net), per phantom of `specs`. Per phantom: a
    Monte Carlo reference accumulated in runs of 4e9 decays until its Poisson floor is <= 1 %
    (or max_ref_runs runs); N5 from the reference; Monte Carlo at N5 and the surrogate each timed
    over `steps` runs (median) after `warmup` untimed runs, wall clock from the run call to the
    estimate on the host. Returns one dict per phantom."""
    import mcpet
    from mcpet import surrogate as S
    names = S._materials()
    rows = []
    log = log or (lambda m: print(m, file=sys.stderr, flush=True))
    for i, sp in enumerate(specs):
        ph = mcpet.phantom_from_arrays(*sp)
        tab = mcpet.load_materials(names, ph.density_max())
        acc = np.zeros(S.GRID)
        ref_decays = 0
        with mcpet.Simulation(S.config(names, 4e9), ph, tab) as sim:
            for j in range(max_ref_runs):
                sim.run(seed=9000 + 17 * i + j)
                acc += sim.rebin("scatter", S.GRID)
                ref_decays += sim.stats()["decays"]
                if S.poisson_floor(acc) <= 0.01:
                    break
        ref = acc / ref_decays
        floor = S.poisson_floor(acc)
        n5 = S.decays_for(ref)
        log(f"time to 5 %: phantom {i}: reference {ref_decays:.3g} decays, floor {floor:.4f}, N5 {n5:.3g}")
        with mcpet.Simulation(S.config(names, n5), ph, tab) as sim:
            for w in range(warmup):
                S.monte_carlo(sim, seed=100 + w)
            mc_t, mc_e = [], []
            for s_ in range(steps):
                e, d, dt, _ = S.monte_carlo(sim, seed=200 + s_)
                mc_t.append(dt)
                mc_e.append(S.nrmse(e, ref))
        with mcpet.Simulation(S.config(names, S.N_LOW), ph, tab) as sim:
            # the phantom's maps are computed once per context (then kept): their first
            # computation is charged to every timed estimate
            t_proj = 0.0
            if est.in_channels == 4:
                t0 = time.perf_counter()
                sim.project(S.GRID, fetch=False)
                t_proj = time.perf_counter() - t0
            for w in range(warmup):
                est.estimate(sim, seed=300 + w)
            su_t, su_e, su_z = [], [], []
            for s_ in range(steps):
                e, info = est.estimate(sim, seed=400 + s_)
                su_t.append(info["seconds"] + t_proj)
                su_e.append(S.nrmse(e, ref))
                su_z.append(info["z"])
        ph.close()
        tab.close()
        t_mc, t_su = float(np.median(mc_t)), float(np.median(su_t))
        ok = [(err <= S.TARGET and z <= est.z_max) for err, z in zip(su_e, su_z)]
        fallback = sum(1 for o in ok if not o) / len(ok)
        t_eff = t_su + fallback * t_mc
        rows.append({"phantom": i, "reference_decays": ref_decays, "reference_floor": floor, "n5_decays": n5,
                     "mc_seconds": t_mc, "mc_nrmse": float(np.median(mc_e)), "surrogate_seconds": t_su,
                     "surrogate_nrmse": float(np.median(su_e)), "surrogate_nrmse_max": float(np.max(su_e)),
                     "consistency_z": float(np.median(su_z)), "fallback_fraction": fallback, "projection_seconds": t_proj,
                     "effective_seconds": t_eff, "speedup": t_mc / t_eff})
        log(f"time to 5 %: phantom {i}: MC {t_mc * 1e3:.1f} ms (nrmse {np.median(mc_e):.4f}), surrogate "
            f"{t_su * 1e3:.2f} ms (nrmse {np.median(su_e):.4f}, z {np.median(su_z):.2f}), speedup {t_mc / t_eff:.1f}")
    return rows


def unet_forward_roofline(est):
    """The U-Net forward of one sample against the dense BF16 MFMA peak: HIP events around the
    network on its stream (mean of 20 forwards after one untimed)."""
    from mcpet import surrogate as S
    x = np.random.default_rng(0).random((1, est.in_channels) + S.GRID).astype(np.float32)
    est.net.forward(x)
    t0 = time.perf_counter()
    dev_ms = []
    for _ in range(20):
        est.net.forward(x)
        dev_ms.append(est.net.forward_ms())
    fwd_host_s = (time.perf_counter() - t0) / 20
    fwd_s = float(np.mean(dev_ms)) / 1e3
    gflop = est.net.gflop
    return {"bound": "mfma", "achieved": gflop / fwd_s / 1e3, "peak": 2500.0, "unit": "TFLOP/s",
            "frac": gflop / fwd_s / 1e3 / 2500.0,
            "traffic": load_traffic("surrogate_unet_forward_traffic.json", "hbm_bytes_per_forward"),
            "kernel": "U-Net forward (one sample; HIP events on the network's stream)", "gflop": gflop,
            "ms": fwd_s * 1e3, "host_ms": fwd_host_s * 1e3}


def bench_surrogate(a):
    """Scatter surrogate workload: time to a 5 %-NRMSE scatter sinogram (time_to_nrmse) on
    a.eval_phantoms evaluation phantoms of one family, drawn with a seed disjoint from the
    training phantoms, and the U-Net forward's roofline."""
    from mcpet import surrogate as S
    weights = a.surrogate_weights or S.WEIGHTS
    if not os.path.exists(weights):
        raise SystemExit(f"bench.py: surrogate weights {weights} missing (train with python -m mcpet.surrogate)")
    specs = eval_phantoms(a.eval_family, a.eval_phantoms, a.eval_seed)
    est = S.ScatterEstimator(weights, width=None)
    rows = time_to_nrmse(specs, est, a.steps, a.warmup)
    roof = unet_forward_roofline(est)
    est.close()
    tot_mc = sum(r["mc_seconds"] for r in rows)
    tot_eff = sum(r["effective_seconds"] for r in rows)
    out = {"metric": "time to a 5%-NRMSE scatter sinogram, Monte Carlo / surrogate", "value": tot_mc / tot_eff,
           "unit": "x (speedup)", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
           "ms_per_step": 1e3 * tot_eff / len(rows), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "bf16 MFMA (fp32 accumulate) + f32 transport", "data": "synthetic",
           "config": {"workload": f"{len(rows)} evaluation phantoms ({EVAL_FAMILIES[a.eval_family]}, seed "
                                  f"{a.eval_seed}), scatter on the 32x56x56 rebinned grid, surrogate input "
                                  f"{S.N_LOW:.0e} decays + activity/attenuation maps, U-Net width {est.width}",
                      "parallelism": "single GPU"},
           "mc_seconds_mean": tot_mc / len(rows), "surrogate_seconds_mean": tot_eff / len(rows),
           "surrogate_nrmse_max": max(r["surrogate_nrmse_max"] for r in rows),
           "fallback_fraction": float(np.mean([r["fallback_fraction"] for r in rows])),
           "roofline": roof, "phantoms": rows}
    print(json.dumps(out))


def time_to_5pct_object(a):
    """The second half of the headline metric beside the default line: time to a 5 %-NRMSE
    scatter sinogram, Monte Carlo against the surrogate, on one evaluation phantom of each
    family (scatter, head, batch; seed disjoint from training), with the U-Net roofline."""
    from mcpet import surrogate as S
    specs = [eval_phantoms(f, 1, a.eval_seed + k)[0] for k, f in enumerate(("scatter", "head", "batch"))]
    est = S.ScatterEstimator(a.surrogate_weights or S.WEIGHTS, width=None)
    steps = max(1, min(a.steps, 5))
    rows = time_to_nrmse(specs, est, steps, 1, max_ref_runs=5)
    roof = unet_forward_roofline(est)
    width = est.width
    est.close()
    tot_mc = sum(r["mc_seconds"] for r in rows)
    tot_eff = sum(r["effective_seconds"] for r in rows)
    return {"metric": "time to a 5%-NRMSE scatter sinogram, Monte Carlo / surrogate", "value": tot_mc / tot_eff,
            "unit": "x (speedup)", "higher_is_better": True,
            "workload": f"one evaluation phantom per family (scatter 128x128x64, head 128x128x96, batch random; seeds "
                        f"{a.eval_seed}-{a.eval_seed + 2}), scatter sinogram rebinned to 32x56x56, Monte Carlo "
                        f"reference of up to 2e10 decays, surrogate = {S.N_LOW:.0e}-decay run + maps + U-Net width "
                        f"{width}; medians of {steps} timed runs each",
            "mc_seconds": [r["mc_seconds"] for r in rows], "surrogate_seconds": [r["effective_seconds"] for r in rows],
            "mc_nrmse": [r["mc_nrmse"] for r in rows], "surrogate_nrmse": [r["surrogate_nrmse"] for r in rows],
            "reference_floor": [r["reference_floor"] for r in rows],
            "fallback_fraction": float(np.mean([r["fallback_fraction"] for r in rows])), "roofline": roof}


def bench_unet(a):
    """The surrogate's U-Net alone on the 32 x 56 x 56 grid (width 32, 4 input channels,
    random inputs): a.steps forward passes of one sample and a.steps training steps of a
    batch of 8, after a.warmup of each; GFLOP counted from the layer shapes."""
    import mcpet
    from mcpet import surrogate as S
    rng = np.random.default_rng(0)
    x = rng.random((8, S.IN_CHANNELS) + S.GRID).astype(np.float32)
    t = rng.random((8,) + S.GRID).astype(np.float32)
    with mcpet.Surrogate(S.GRID, in_channels=S.IN_CHANNELS, width=S.WIDTH, seed=1) as net:
        for _ in range(a.warmup):
            net.forward(x[:1])
            net.train_step(x, t, lr=1e-4)
        t0 = time.perf_counter()
        dev_ms = []
        for _ in range(a.steps):
            net.forward(x[:1])
            dev_ms.append(net.forward_ms())
        fwd_host = (time.perf_counter() - t0) / a.steps
        fwd = float(np.mean(dev_ms)) / 1e3       # HIP events around the network on its stream
        t0 = time.perf_counter()
        for _ in range(a.steps):
            net.train_step(x, t, lr=1e-4)
        trn = (time.perf_counter() - t0) / a.steps
        gflop = net.gflop
        n_params = net.n_params
    tf = gflop / fwd / 1e3
    out = {"metric": "U-Net forward throughput", "value": tf, "unit": "TFLOP/s", "n_gpus": 1, "steps": a.steps,
           "warmup": a.warmup, "ms_per_step": fwd * 1e3, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "bf16 MFMA, fp32 accumulate", "data": "synthetic",
           "config": {"workload": f"3D U-Net width {S.WIDTH}, {n_params} parameters, grid {S.GRID}, forward of one sample; "
                                  "training step of 8 samples (forward, backward, Adam)", "parallelism": "single GPU"},
           "forward_gflop": gflop, "forward_host_ms": fwd_host * 1e3, "train_step_ms": trn * 1e3,
           "train_tflops": 3 * 8 * gflop / trn / 1e3,
           "roofline": {"bound": "mfma", "achieved": tf, "peak": 2500.0, "unit": "TFLOP/s", "frac": tf / 2500.0,
                        "traffic": load_traffic("surrogate_unet_forward_traffic.json", "hbm_bytes_per_forward"),
                        "kernel": "U-Net forward (HIP events on the network's stream)"}}
    print(json.dumps(out))


def bench_serve(a):
    """Scatter estimates served to Poisson-arriving requests: a.serve_requests phantoms cycling
    through the scatter, head and batch families (mixed sizes, seeds disjoint from training),
    arriving at a.serve_rate per second, each answered by the serving loop (low-count run,
    rebinning and maps on the device, the U-Net, the consistency check, Monte Carlo fallback,
    every a.spot_every-th request spot-checked against Monte Carlo at N5) under an HBM budget.
    Latency = completion - arrival; the phantoms are built before the clock starts."""
    import mcpet
    from mcpet import surrogate as S
    fams = ("scatter", "head", "batch")
    n = a.serve_requests
    per = {f: S.family_phantoms(f, (n + 2) // 3, a.eval_seed + 11 * k) for k, f in enumerate(fams)}
    specs = [per[fams[i % 3]][i // 3] for i in range(n)]
    phs = [mcpet.phantom_from_arrays(*sp) for sp in specs]
    names = S._materials()
    tab = mcpet.load_materials(names, np.max([p.density_max() for p in phs], axis=0))
    est = S.ScatterEstimator(a.surrogate_weights or S.WEIGHTS, width=None)
    # warm-up: one request alone (kernels loaded, per-device buffers and crystal-pair lists built)
    svc = S.ScatterService(tab, est, budget_bytes=a.serve_budget_gb << 30, max_running=16)
    svc.submit(phs[0], seed=1)
    svc.drain()
    svc.close()
    svc = S.ScatterService(tab, est, budget_bytes=a.serve_budget_gb << 30, max_running=16,
                           streams_per_step=a.serve_chunk, spot_check_every=a.spot_every)
    rng = np.random.default_rng(a.eval_seed)
    arrival = np.cumsum(rng.exponential(1.0 / a.serve_rate, n))
    sid_of = {}
    i = 0
    t0 = time.perf_counter()
    while i < n or svc.by_rid:
        now = time.perf_counter() - t0
        while i < n and arrival[i] <= now:
            sid_of[svc.submit(phs[i], seed=1000 + i)] = i
            i += 1
        if svc.by_rid:
            svc.step()
        elif i < n:
            time.sleep(max(0.0, arrival[i] - (time.perf_counter() - t0)))
    makespan = time.perf_counter() - t0
    lat, kinds = [], []
    for sid, k in sid_of.items():
        job = svc.jobs[sid]
        done = job["t0"] + job["seconds"] - t0
        lat.append(1e3 * (done - arrival[k]))
        kinds.append(job["stage"])
    st = svc.stats()
    svc.close()
    est.close()
    for p in phs:
        p.close()
    tab.close()
    lat = np.array(lat)
    surrogate_lat = lat[[k == "surrogate" for k in kinds]]
    out = {"metric": "scatter-estimate requests served per second (Poisson arrivals)", "value": n / makespan,
           "unit": "requests/s", "n_gpus": 1, "steps": n, "warmup": 1, "ms_per_step": 1e3 * makespan / n,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "bf16 MFMA (fp32 accumulate) + f32 transport", "data": "synthetic",
           "config": {"workload": f"{n} requests, Poisson arrivals at {a.serve_rate:g}/s, phantoms cycling through the "
                                  f"scatter / head / batch families (seed {a.eval_seed}), HBM budget "
                                  f"{a.serve_budget_gb} GiB, 16 running, {a.serve_chunk} streams per request and step, spot "
                                  f"check every {a.spot_every}",
                      "parallelism": "single GPU"},
           "latency_p50_ms": float(np.percentile(lat, 50)), "latency_p99_ms": float(np.percentile(lat, 99)),
           "latency_max_ms": float(lat.max()),
           "surrogate_latency_p50_ms": float(np.percentile(surrogate_lat, 50)) if len(surrogate_lat) else None,
           "surrogate_latency_p99_ms": float(np.percentile(surrogate_lat, 99)) if len(surrogate_lat) else None,
           "answered_by": {k: kinds.count(k) for k in sorted(set(kinds))}, "fallbacks": st["fallbacks"],
           "spot_check_errors": st["spot_checks"], "surrogate_disabled": st["disabled"],
           "peak_admitted_bytes": st["server"].get("peak_admitted_bytes")}
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--decays", type=float, default=0.0, help="expected decays (0: 1e9 head, 5e9 torso)")
    ap.add_argument("--tpb", type=int, default=32, help="decay streams per emitting voxel")
    ap.add_argument("--kernel", choices=["persistent", "plain"], default="persistent")
    ap.add_argument("--blocks-per-cu", type=int, default=0)
    ap.add_argument("--block-size", type=int, default=0)
    ap.add_argument("--phase-policy", type=int, default=0, help="0 auto, 1 most lanes, 2 least waste")
    ap.add_argument("--phase-cost", type=str, default="", help="comma-separated relative phase costs (DECAY, SOURCE2, STEP, INTERACT, TALLY, CTAU, CPZ)")
    ap.add_argument("--voxel-format", type=int, default=0, help="0 auto (palette indices), 1 8-byte voxel words")
    ap.add_argument("--workload", choices=["head", "torso", "batch", "surrogate", "unet", "serve"], default="head",
                    help="head: the 1e9-decay head phantom; torso: the 5e9-decay torso phantom with beta+ "
                         "branching, prompt gammas and dose tallies; batch: 48 varied phantoms in one launch; "
                         "surrogate: time to a 5%%-NRMSE scatter sinogram, Monte Carlo vs the learned surrogate; "
                         "unet: the surrogate network alone; serve: scatter estimates for Poisson-arriving requests")
    ap.add_argument("--eval-phantoms", type=int, default=6)
    ap.add_argument("--eval-seed", type=int, default=777)
    ap.add_argument("--eval-family", choices=sorted(EVAL_FAMILIES), default="scatter",
                    help="surrogate workload: the phantom family of the evaluation")
    ap.add_argument("--surrogate-weights", default="", help="U-Net parameters (.npy); default: the packaged ones")
    ap.add_argument("--batch-size", type=int, default=48)
    ap.add_argument("--serve-requests", type=int, default=60)
    ap.add_argument("--serve-rate", type=float, default=20.0, help="serve workload: mean arrivals per second")
    ap.add_argument("--serve-budget-gb", type=int, default=32)
    ap.add_argument("--serve-chunk", type=int, default=1 << 20,
                    help="serve workload: decay streams per request and serving-loop step")
    ap.add_argument("--spot-every", type=int, default=20, help="serve workload: spot-check every k-th request")
    ap.add_argument("--singles", action="store_true",
                    help="torso workload: the singles pipeline (time sort, 4.5 ns coincidence window, randoms) "
                         "instead of the fused pair tallies")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="CPU baseline threads (0: the CPUs this process may use, capped by OMP_NUM_THREADS)")
    ap.add_argument("--no-surrogate", action="store_true",
                    help="head workload: leave out the time-to-5%%-NRMSE object")
    ap.add_argument("--cpu-stride", type=int, default=0, help="CPU baseline voxel stride (0: 5 head, 50 torso)")
    ap.add_argument("--copy-calibration", type=float, default=0,
                    help="only run the library's device copy of this many bytes (HBM counter calibration) and exit")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build measures one MI355X only; multi-GPU runs are not supported "
              "(decays, phantoms and requests would shard by stream range, phantom and request)", file=sys.stderr)
        sys.exit(2)

    import mcpet
    if a.copy_calibration:
        n = int(a.copy_calibration)
        ms = mcpet.selftest_copy(n, iters=5)
        print(json.dumps({"copy_bytes": n, "iters": 5, "ms_per_copy": ms, "GB/s": 2 * n / ms / 1e6}))
        return
    from mcpet import phantoms
    if a.workload == "batch":
        return bench_batch(a)
    if a.workload == "surrogate":
        return bench_surrogate(a)
    if a.workload == "unet":
        return bench_unet(a)
    if a.workload == "serve":
        return bench_serve(a)
    torso = a.workload == "torso"
    n_decays = a.decays or (5.0e9 if torso else 1.0e9)
    stride = a.cpu_stride or (50 if torso else 5)
    nd = f"{n_decays:.0e}".replace("e+0", "e").replace("e+", "e")
    if torso:
        # 50 % beta+ and a 603 keV prompt gamma per decay (the interaction tables continued past
        # the material files' 515 keV), 64-bit dose tallies over every voxel, a 46 cm detector
        # around the 64 x 64 cm box
        mat, dens, act, vs = phantoms.torso(n_decays=n_decays)
        names = phantoms.torso_materials(MATERIALS)
        nz, ny, nx = mat.shape
        # singles pipeline: about 0.35 detected singles per decay; room for 0.43 (the list's
        # 32-bit indices allow 2^31 - 1 records)
        singles_kw = dict(tally_psf_sinogram=3, singles_capacity=int(min(2 ** 31 - 1, 0.5 * n_decays + (1 << 20))),
                          coincidence_window_ps=4500) if a.singles else {}
        cfg = phantoms.make_config(seed=20261202, threads_per_block=1, acquisition_time_s=10.0, nrad=168, radius=-46.0,
                                   materials=names, positron_fraction=0.5, prompt_gamma_fraction=1.0,
                                   prompt_gamma_energy=603.0e3, material_dose=True, voxel_dose=True,
                                   roi=(0, nx - 1, 0, ny - 1, 0, nz - 1),
                                   stream_decays=int(max(4, np.ceil(n_decays / 2.0 ** 20))), **singles_kw)
        workload = (f"torso phantom 256x256x200 at 2.5 mm, 7 tissues, 50% beta+ and a 603 keV prompt gamma per decay, "
                    f"{nd} decays in 10 s, sinogram 168x168x1293, 64-bit voxel and material dose, " +
                    ("singles list, time sort and 4.5 ns coincidence window (trues, scatter, randoms, spurious)"
                     if a.singles else "fused transport + tallies"))
    else:
        mat, dens, act, vs = phantoms.head(n_decays=n_decays)
        names = phantoms.head_materials(MATERIALS)
        cfg = phantoms.make_config(seed=20261202, threads_per_block=a.tpb, acquisition_time_s=10.0, nrad=168,
                                   materials=names)
        workload = (f"head phantom 128x128x96 at 2 mm, 8 tissues, 18F, {nd} decays in 10 s, "
                    f"sinogram 168x168x1293, fused transport + tallies")
    cfg.voxel_format = a.voxel_format
    ph = mcpet.phantom_from_arrays(mat, dens, act, vs)
    e_max = 610.0e3 if torso else 0.0
    tab = mcpet.load_materials(names, ph.density_max(), e_max=e_max)
    kernel = mcpet.KERNEL_PERSISTENT if a.kernel == "persistent" else mcpet.KERNEL_PLAIN
    sim = mcpet.Simulation(cfg, ph, tab)
    run_kw = dict(kernel=kernel, block_size=a.block_size, blocks_per_cu=a.blocks_per_cu, phase_policy=a.phase_policy,
                  phase_cost=[int(x) for x in a.phase_cost.split(",")] if a.phase_cost else None)
    for w in range(a.warmup):
        sim.run(seed=1000 + w, **run_kw)
        sim.sync()
    sim.sync()
    coinc = decays = 0
    kernel_ms = []
    per_step = []
    singles = []
    t0 = time.perf_counter()
    for s in range(a.steps):
        sim.run(seed=2000 + s, **run_kw)
        st = sim.stats()               # synchronises; decays and coincidences of this step
        coinc += st["coincidences"]
        decays += st["decays"]
        kernel_ms.append(st["kernel_ms"])
        per_step.append((st["decays"], st["coincidences"]))
        if a.singles:
            cc = sim.fetch(sinograms=False, images=False, spectrum=False, psf=False, dose=False,
                           singles=False)["coincidences"]
            singles.append((st["singles"], st["coinc_ms"], cc))
    sim.sync()
    elapsed = time.perf_counter() - t0
    # Algorithmic event counts of the timed steps: the same seeds replayed (untimed) with the
    # instrumented kernel, which performs the identical simulation and counts every event.
    agg = None
    for s in range(a.steps):
        sim.run(seed=2000 + s, count_events=True, **run_kw)
        st = sim.stats()
        if (st["decays"], st["coincidences"]) != per_step[s]:
            raise SystemExit(f"bench.py: instrumented replay of step {s} differs from the timed run")
        agg = st if agg is None else {k: (agg[k] + st[k]) if not isinstance(st[k], list) else
                                      [x + y for x, y in zip(agg[k], st[k])] for k in agg}
    sim.close()

    ops, byts = algorithmic(agg, voxel_dose=torso)
    kt = sum(kernel_ms) / 1e3
    avg_ms = sum(kernel_ms) / len(kernel_ms)
    valu = ops / kt
    hbm = byts / kt
    # measured HBM bytes per launch of this workload at its default size (a profile of that run)
    traffic = None
    if not a.decays and not a.singles:
        traffic = load_traffic("transport_torso_traffic.json" if torso else "transport_head_traffic.json")
    out = {
        "metric": "detected coincidences/s on 1 MI355X", "value": coinc / elapsed, "unit": "coincidences/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": 1e3 * elapsed / a.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": workload,
                   "decays_per_step": decays / a.steps, "threads_per_block": 1 if torso else a.tpb, "kernel": a.kernel,
                   "parallelism": "single GPU"},
        "decays_per_s": decays / elapsed, "photon_histories_per_s": 2 * decays / elapsed,
        "kernel_ms_avg": avg_ms,
        # The transport uses no MFMA: its roofline is HBM, algorithmic bytes per launch over the
        # kernel time. Its binding resource is VALU issue with dependent gathers; that figure
        # (algorithmic lane operations over the non-packed VALU peak) is reported beside it.
        "roofline": {"bound": "hbm", "achieved": hbm / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": hbm / HBM_PEAK, "traffic": traffic, "algorithmic_bytes_per_launch": byts / a.steps,
                     "kernel": "k_transport_sm (persistent)",
                     "valu": {"achieved": valu / 1e12, "peak": VALU_PEAK / 1e12, "unit": "Tlane-op/s",
                              "frac": valu / VALU_PEAK, "algorithmic_lane_ops_per_launch": ops / a.steps}},
        "events_per_decay": {k: agg[k] / max(1, decays) for k in ("woodcock_steps", "material_changes", "compton",
                                                                   "rayleigh", "photoelectric", "coincidences",
                                                                   "rn_draws")},
    }
    if any(agg["phase_exec"]):
        tot_cyc = max(1, sum(agg["phase_cycles"]))
        out["phase_stats"] = {n: {"exec_per_decay": e / max(1, decays), "mean_lanes": l / max(1, e),
                                  "cycle_share": c / tot_cyc, "cycles_per_exec": c / max(1, e)}
                              for n, e, l, c in zip(("decay", "source2", "step", "interact", "tally", "ctau", "cpz", "detect", "source3"),
                                                    agg["phase_exec"], agg["phase_lanes"], agg["phase_cycles"])}
    if singles:
        out["singles_pipeline"] = singles_roofline(singles, cfg.acquisition_time_ps)
    if not a.no_cpu_baseline and not a.singles:
        out["cpu_baseline"] = cpu_baseline(cfg, ph, vs, names, baseline_threads(a.cpu_threads), stride, e_max)
    if not torso and not a.no_surrogate:
        out["time_to_5pct_nrmse"] = time_to_5pct_object(a)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
