#!/usr/bin/env python3
"""Throughput of the MI355X phonon engine on psim's scaled steady-state workload.

Workload (one "step" = one complete simulation run: emission of the whole budget, then every
phonon carried through 5000 measurement steps of 2 ps, tallies reduced on the GPU):
the conforming L-shaped silicon structure (40 sensors, 80 triangles, walls 320 / 280 K,
specularity 0.3, T_init 300 K), full (non-deviational) mode, 1e7 phonons, FP64 tallies.

  value      phonon-steps per second of the timed runs (a phonon-step is one live phonon carried
             across one measurement interval: the reference's ++step at
             psim/src/modelSimulator.cpp:182-183), with the model already resident in HBM.
  roofline   the transport kernel against the step-synchronous HBM model of this path: 120 bytes
             per phonon-step (the phonon record read and written once per step) over its
             HIP-event time on the engine stream.  What the counters show limits it (`limiter`,
             VALU issue, and the measured HBM bytes per phonon-step) is copied from profiles/,
             labelled with the engine commit it was measured on and whether that is this tree's.
  cpu_baseline  the reference psim (oracle/_ref/psim_ref, built from the reference sources) on
             a bounded sample of the same workload (1e5 phonons): one process per CPU of the
             job's share (the affinity set, at most 16) started together, their aggregate rate
             (its std::execution::par loop is serial in this image), with one process alone
             beside it and the host's CPU model, CPU count, affinity and cgroup quota;
             phonon-steps are counted by the CPU oracle on the same model.
  time_to_steady_state  on the 1000x100 nm bar (10 sensors, walls 310/290 K, deviational, every
             step recorded, 1e7 phonons, 2000 steps of 5 ps): the step after which the RMS over
             sensors of the change in 100-step window-mean temperature stays below 0.05 K for 3
             windows (psim_amd.analysis), then the same simulation cut at that step (phonon budget
             scaled so the energy per phonon is unchanged) timed on the GPU.  The reference runs the
             cut simulation too, its budget split over one process per CPU of the job's share
             started together (cpu_measured); its one-process time is extrapolated from its rate on
             the same bar at 1e5 phonons (cpu_extrapolated).

With --workload film it measures the transient case instead: the 1500x150 nm film (30 sensors,
baseline 300 K, a 20 K burst for 100 ps into sensors 0-2), 1e8 phonons, 20000 steps of 1 ps, every
step tallied (snapshots every 100 steps), one run after no warmup; its CPU baseline is the
reference on the same film at 1e4 phonons.

With --workload batch it measures batched geometries: 256 independent structures from
psim_amd.configs.batch_geometries(seed 0) (straight or kinked Si bars, 4-200 sensors, walls
U(250, 350) K, specularity U(0, 1), 1e4-1e6 phonons each, deviational steady state, 2000 steps of
5 ps) run as one engine upload and one run (psim_batch_run).  `value` is their phonon-steps per
second of GPU run time; the same models run one after the other on the GPU, and the reference on
a sample of them (scaled to 1e4 phonons each), are reported beside it.

With --workload surrogate it measures the warm start: held-out structures
(psim_amd.train_surrogate.structures) run from a cold start and from the trained surrogate's
prediction (8e6 phonons, 5000 steps of 5 ps, periodic mode); `value` is the ratio of their total
times to steady state in simulated time (RMS over sensors of the change in 100-step window-mean T
below 0.005 x |t_hot - t_cold| for 3 windows); cold_over_warm_gpu_seconds is the same ratio in GPU
time, from runs cut at their measured steady state.  The roofline is the surrogate's fused BF16 MFMA forward product
(gemm_nt_kernel, 65536 x 512 x 512) against the dense bf16 peak, timed with HIP events; its CPU
baseline is the reference's rate on two of the held-out structures.

With --workload gemm it measures the surrogate's products alone: the network's fused products of a
hidden layer at 4096 and 65536 rows (psim_mlp_time_product), hipBLASLt's plain product beside them
and a training step at batch 4096.

Usage: python bench.py [--gpus 1] [--steps 3] [--warmup 1] [--workload lshape|film|batch|surrogate|gemm]
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

BYTES_PER_PHONON_STEP = 120.0
HBM_PEAK_GBPS = 8000.0
ENGINE_SOURCES = ("psim_amd/csrc/engine.hip", "psim_amd/csrc/engine_device.hpp", "psim_amd/csrc/psim_rng.h",
                  "psim_amd/csrc/psim_math.h", "include/psim_gpu.h")


def engine_sha256() -> str:
    """Hash of the transport engine's sources (the code a counter profile describes)."""
    import hashlib

    h = hashlib.sha256()
    for rel in ENGINE_SOURCES:
        with open(os.path.join(ROOT, rel), "rb") as f:
            h.update(f.read())
    return h.hexdigest()


def profiled_counters():
    """The transport kernel's counter figures on the L-shape (profiles/), with `current` telling
    whether they were measured on the engine sources of this tree."""
    path = os.path.join(ROOT, "profiles", "transport_kernel_lshape_1e7_counters.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        c = json.load(f)
    c["current"] = c.get("engine_sha256") == engine_sha256()
    return c


def host_cpus() -> dict:
    """What this host gives the CPU baseline: CPU model, the CPUs the OS reports and the ones this
    process may run on, the cgroup's CPU quota, and the processes the baseline starts (one per
    CPU of the job's share: the affinity set, at most 16, the CPU share a one-GPU job is given on
    the GPU hosts this was measured on)."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, period = f.read().split()[:2]
            quota = None if q == "max" else float(q) / float(period)
    except (OSError, ValueError):
        pass
    affinity = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    procs = max(1, min(16, affinity, int(quota) if quota else affinity))
    return {"cpu_model": model, "nproc": os.cpu_count(), "affinity": affinity, "cgroup_cpu_quota": quota,
            "processes": procs}


def reference_rate_all_cores(cfg: dict, procs: int):
    """(aggregate phonon-steps/s, group wall s, processes) of `procs` independent reference
    processes on cfg started together, one per core: the reference's std::execution::par loop
    runs serially in this image (no TBB headers, so libstdc++ uses its serial backend), and
    independent processes are the most the host's cores give it.  None without the reference
    build."""
    import oracle
    from oracle import refmodel

    ref_bin = os.path.join(ROOT, "oracle", "_ref", "psim_ref")
    if not os.path.exists(ref_bin):
        return None
    counted = oracle.Model(cfg).run(0, nthreads=procs).counters["phonon_steps"]
    with tempfile.TemporaryDirectory() as wd:
        txt = os.path.join(wd, "model.txt")
        with open(txt, "w") as f:
            f.write(refmodel.json_to_text(cfg))
        t0 = time.perf_counter()
        running = [subprocess.Popen([ref_bin, txt, os.path.join(wd, f"out{i}.json")], stdout=subprocess.DEVNULL)
                   for i in range(procs)]
        for p in running:
            if p.wait(timeout=900) != 0:
                raise RuntimeError("reference run failed")
        wall = time.perf_counter() - t0
    return counted * procs / wall, wall, procs


def reference_group_wall(cfg: dict, procs: int) -> float:
    """Wall seconds of `procs` reference processes on cfg started together (until the last ends)."""
    from oracle import refmodel

    ref_bin = os.path.join(ROOT, "oracle", "_ref", "psim_ref")
    with tempfile.TemporaryDirectory() as wd:
        txt = os.path.join(wd, "model.txt")
        with open(txt, "w") as f:
            f.write(refmodel.json_to_text(cfg))
        t0 = time.perf_counter()
        running = [subprocess.Popen([ref_bin, txt, os.path.join(wd, f"out{i}.json")], stdout=subprocess.DEVNULL)
                   for i in range(procs)]
        for p in running:
            if p.wait(timeout=900) != 0:
                raise RuntimeError("reference run failed")
        return time.perf_counter() - t0


def reference_rate(cfg: dict) -> tuple:
    """(phonon-steps/s, wall s, phonon-steps, kind) of one single-threaded reference run of cfg (the
    oracle, single-threaded, where the reference build is absent); phonon-steps counted by the
    oracle on the same model."""
    import oracle
    from oracle import refmodel

    threads = host_cpus()["processes"]
    counted = oracle.Model(cfg).run(0, nthreads=threads).counters["phonon_steps"]
    ref_bin = os.path.join(ROOT, "oracle", "_ref", "psim_ref")
    if os.path.exists(ref_bin):
        with tempfile.TemporaryDirectory() as wd:
            txt = os.path.join(wd, "model.txt")
            out = os.path.join(wd, "out.json")
            with open(txt, "w") as f:
                f.write(refmodel.json_to_text(cfg))
            subprocess.run([ref_bin, txt, out], check=True, stdout=subprocess.DEVNULL, timeout=600)
            with open(out) as f:
                wall = json.load(f)["wall_s"]
        return counted / wall, wall, counted, "reference"
    t0 = time.perf_counter()
    oracle.Model(cfg).run(0, nthreads=1)
    wall = time.perf_counter() - t0
    return counted / wall, wall, counted, "port"


def workload(name: str, phonons: int) -> tuple:
    """(model dict, config description) of a benchmark workload."""
    from psim_amd import configs

    if name == "film":
        cfg = configs.film_transient(num_phonons=phonons)
        desc = {"workload": "1500x150 nm Si film (30 sensors, 60 triangles), transient deviational, baseline "
                            "300 K, 20 K burst for 100 ps into sensors 0-2, snapshots every 100 steps",
                "phonons": phonons, "measurement_steps": 20000, "dt_ps": 1.0}
    else:
        cfg = configs.lshape_full(num_phonons=phonons)
        desc = {"workload": "L-shaped Si (40 sensors, 80 triangles), full-mode steady state, walls 320/280 K, "
                            "specularity 0.3", "phonons": phonons, "measurement_steps": 5000, "dt_ps": 2.0}
    desc.update({"tallies": "fp64", "parallelism": "single GPU"})
    return cfg, desc


def cpu_baseline(name: str, sample_phonons: int) -> dict:
    """The reference on a sample of the benchmark workload: one process per CPU of the job's share
    started together (host_cpus), and one process alone."""
    cfg, desc = workload(name, sample_phonons)
    host = host_cpus()
    rate, wall, counted, kind = reference_rate(cfg)
    sample = (f"{desc['workload']}, {sample_phonons} phonons x {desc['measurement_steps']} steps of "
              f"{desc['dt_ps']} ps; {counted} phonon-steps per run counted by the oracle on the same model")
    single = {"value": rate, "cores": 1, "wall_s": wall}
    multi = reference_rate_all_cores(cfg, host["processes"]) if kind == "reference" else None
    if multi is None:
        return {"value": rate, "unit": "phonon-steps/s", "cores": 1, "kind": kind,
                "sample": sample + ", one process", "wall_s": wall, "host": host}
    m_rate, m_wall, n = multi
    return {"value": m_rate, "unit": "phonon-steps/s", "cores": n, "kind": kind,
            "sample": sample + f", {n} independent processes started together (one per CPU of the job's share)",
            "wall_s": m_wall, "one_process": single, "scaling_over_one_process": m_rate / rate, "host": host}


def steady_state(engine, phonons: int, cpu: bool) -> dict:
    """Time to steady state of the deviational bar on the GPU, and of the reference on the same cut run."""
    import psim_amd
    from psim_amd import analysis, configs

    n_steps, sim_ns = 2000, 10.0
    full = psim_amd.Model(configs.linear_bar(num_phonons=phonons, sim_type=1, step_interval=100).to_dict())
    full.run(engine, 0)
    steps = analysis.time_to_steady_state(full.results(0).temps)
    out = {"workload": f"1000x100 nm Si bar, 10 sensors, walls 310/290 K, deviational (t_eq 300 K), every step "
                       f"recorded, {phonons} phonons, {n_steps} steps of 5 ps",
           "criterion": "RMS over sensors of the change in 100-step window-mean T < 0.05 K for 3 windows",
           "steps": steps}
    if steps is None:
        return out
    frac = steps / n_steps
    cut = configs.linear_bar(num_phonons=int(round(phonons * frac)), num_measurements=steps, sim_time=sim_ns * frac,
                             sim_type=1, step_interval=100).to_dict()
    m = psim_amd.Model(cut)
    m.run(engine, 1)
    m.run(engine, 0)
    gpu_s = engine.last_timing()["run_ms"] / 1e3
    ps = m.results(0).counters["phonon_steps"]
    out.update({"sim_ns": sim_ns * frac, "gpu_wall_s": gpu_s, "phonon_steps": ps})
    if cpu:
        # Not a run: the reference's time for these phonon-steps at the rate it reaches on the same
        # bar at 1e5 phonons (one process, and the job's share of CPUs as independent processes).
        bar = configs.linear_bar(num_phonons=100_000, sim_type=1, step_interval=100).to_dict()
        rate, _, _, kind = reference_rate(bar)
        out["cpu_extrapolated"] = {
            "method": "phonon-steps of the cut GPU run divided by the reference's measured rate on the same bar "
                      "at 1e5 phonons; the reference run itself is not timed (about 3 core-minutes)",
            "kind": kind, "one_process": {"rate": rate, "wall_s": ps / rate, "speedup": ps / rate / gpu_s}}
        multi = reference_rate_all_cores(bar, host_cpus()["processes"]) if kind == "reference" else None
        if multi is not None:
            m_rate, _, n = multi
            out["cpu_extrapolated"]["all_processes"] = {"processes": n, "rate": m_rate, "wall_s": ps / m_rate,
                                                       "speedup": ps / m_rate / gpu_s}
            # Measured: the cut run itself on the reference, its phonon budget split over one process
            # per CPU of the job's share, all started together (each process a full simulation of
            # the same bar and steps with 1/n of the phonons; their mean is the same estimate).
            part = dict(cut, settings=dict(cut["settings"], num_phonons=max(1, cut["settings"]["num_phonons"] // n)))
            wall = reference_group_wall(part, n)
            out["cpu_measured"] = {"processes": n, "wall_s": wall, "speedup": wall / gpu_s,
                                   "method": f"the cut run ({cut['settings']['num_phonons']} phonons, {steps} steps) "
                                             f"on the reference build as {n} processes of "
                                             f"{part['settings']['num_phonons']} phonons each, started together; "
                                             "wall time until the last one ends"}
    return out


def build_models(cfgs: list) -> list:
    """psim_amd.Model of every config, built on host threads (ctypes releases the GIL)."""
    import psim_amd
    from concurrent.futures import ThreadPoolExecutor

    with ThreadPoolExecutor(max_workers=16) as ex:
        return list(ex.map(psim_amd.Model, cfgs))


def batch_bench(args) -> dict:
    import psim_amd
    from psim_amd import configs

    cfgs = configs.batch_geometries(seed=0, count=256)
    engine = psim_amd.Engine(0)
    t0 = time.perf_counter()
    models = build_models(cfgs)
    t_build = time.perf_counter() - t0
    for w in range(max(1, args.warmup)):
        psim_amd.run_batch(engine, models, seed=10_000 + w)
    # launch shape for this batch (the batch is now uploaded); no choice changes any result
    tune = None if args.no_autotune else engine.autotune(seed=20_000)
    gpu_s = wall = 0.0
    steps_total = events_total = 0
    for k in range(args.steps):
        t0 = time.perf_counter()
        counters = psim_amd.run_batch(engine, models, seed=k)
        wall += time.perf_counter() - t0
        gpu_s += engine.last_timing()["run_ms"] / 1e3
        steps_total += counters["phonon_steps"]
        # events of the event loop: measurement boundaries passed, scatters and wall impacts
        events_total += counters["phonon_steps"] + counters["scatters"] + counters["impacts"]
    value = steps_total / gpu_s
    # the uploaded batch replayed from a captured graph (capture on the first call)
    n_sensors = sum(len(c["sensors"]) for c in cfgs)
    n_record = models[0].num_record
    engine.run_graph(n_sensors, n_record, 30_000)
    g_steps, g_s = 0, 0.0
    for k in range(args.steps):
        _, _, gc = engine.run_graph(n_sensors, n_record, 30_001 + k)
        g_steps += gc["phonon_steps"]
        g_s += engine.last_timing()["run_ms"] / 1e3
    # the same models one at a time (each its own upload and run), with the default launch shape
    engine.set_window(0)
    engine.set_grid(8)
    engine.set_block(256)
    seq_steps, seq_gpu = 0, 0.0
    for m in ([] if args.no_sequential else models):
        m.run(engine, 0)
        seq_steps += m.results(0).counters["phonon_steps"]
        seq_gpu += engine.last_timing()["run_ms"] / 1e3
    line = {
        "metric": "phonon-steps/s", "value": value, "unit": "phonon-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": gpu_s * 1e3 / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "256 independent Si structures (psim_amd.configs.batch_geometries seed 0): straight or "
                               "kinked bars, 4-200 sensors, walls U(250,350) K, specularity U(0,1), 1e4-1e6 "
                               "phonons each, deviational steady state, 2000 steps of 5 ps, one batched run",
                   "models": len(models), "phonons": int(sum(c["settings"]["num_phonons"] for c in cfgs)),
                   "tallies": "fp64", "parallelism": "single GPU"},
        "batch": {"gpu_run_s": gpu_s / args.steps, "end_to_end_s": wall / args.steps,
                  "host_model_build_s": t_build, "phonon_steps": steps_total // args.steps,
                  "events_per_s": events_total / gpu_s,
                  "events_per_phonon_step": events_total / steps_total,
                  "autotune": tune, "graph_replay": {"value": g_steps / g_s, "gpu_run_s": g_s / args.steps}},
        "sequential_gpu": None if args.no_sequential else {
            "value": seq_steps / seq_gpu, "gpu_run_s": seq_gpu, "phonon_steps": seq_steps,
            "batch_speedup": value / (seq_steps / seq_gpu)},
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        import copy
        rate_steps, rate_wall = 0, 0.0
        for cfg in cfgs[:4]:
            c = copy.deepcopy(cfg)
            c["settings"]["num_phonons"] = 10_000
            _, w, n, kind = reference_rate(c)
            rate_steps += n
            rate_wall += w
        line["cpu_baseline"] = {"value": rate_steps / rate_wall, "unit": "phonon-steps/s", "cores": 1, "kind": kind,
                                "sample": "the first 4 structures of the batch at 1e4 phonons each, one process, "
                                          "run one after the other", "wall_s": rate_wall}
    engine.close()
    return line


MFMA_BF16_PEAK_TFLOPS = 2500.0  # MI355X dense bf16
# HBM bytes of one 65536 x 512 x 512 fused forward product, per kernel the autotuner may pick for it,
# copied from the counter profiles (FETCH_SIZE x 1.9910 + WRITE_SIZE x 0.9666, the copy-kernel
# calibration, per dispatch; counters cannot be read in this process):
# - weights resident in LDS, 8 waves: 34992 KiB and 65536 KiB (profiles/gemm_nt_bres_kernel_surrogate.md);
# - the 128 x 128 tile: 39469 KiB and 83280 KiB (profiles/gemm_nt_kernel_surrogate.md).
GEMM_FWD_65536_TRAFFIC = {
    "512x128 (weights resident in LDS)": (136.2e6, "gemm_nt_bres_kernel", "profiles/gemm_nt_bres_kernel_surrogate.md"),
    "128x128": (162.9e6, "gemm_nt_kernel", "profiles/gemm_nt_kernel_surrogate.md"),
}


def hipblaslt_gemm(M: int, N: int, K: int) -> dict:
    """torch.matmul (hipBLASLt) on bf16 [M][K] x [K][N] with a bf16 result, HIP events over 50 products."""
    import torch

    dev = torch.device("cuda:0")
    a = torch.randn(M, K, device=dev, dtype=torch.bfloat16)
    b = torch.randn(K, N, device=dev, dtype=torch.bfloat16)
    for _ in range(3):
        a @ b
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(50):
        a @ b
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / 50
    return {"M": M, "N": N, "K": K, "product": "torch.matmul bf16 (hipBLASLt)", "ms": ms,
            "tflops": 2.0 * M * N * K / (ms / 1e3) / 1e12}


def surrogate_products(rows: int = 0) -> tuple:
    """The network's own fused products of a hidden 512 x 512 layer (psim_mlp_time_product), at the
    training batch (4096 rows), at `rows` rows (one held-out forward pass, when given) and at 65536
    rows: forward (bias, ReLU and bf16 rounding in the epilogue), forward while training (plus the
    transposed copy), weight gradient (on the transposed copies, split over the rows) and input
    gradient (through the ReLU); hipBLASLt on the plain 65536-row product beside them; and one
    training step (forward, backward, Adam) at batch 4096, timed over 50 steps."""
    import numpy as np
    from psim_amd import surrogate

    trainer = surrogate.Surrogate(seed=1)
    shapes = [("forward_hidden_b4096", "forward_train", 4096), ("weight_grad_hidden_b4096", "weight_grad", 4096),
              ("input_grad_hidden_b4096", "input_grad", 4096)]
    if rows:
        shapes.append((f"forward_hidden_{rows}_rows", "forward", rows))
    shapes += [("forward_train_hidden_65536_rows", "forward_train", 65536),
               ("weight_grad_hidden_65536_rows", "weight_grad", 65536),
               ("forward_hidden_65536_rows", "forward", 65536)]
    gemm = {}
    for name, kind, r in shapes:
        ms, tile = trainer.time_product(kind, r, reps=50)
        M, N, K = (512, 512, r) if kind == "weight_grad" else (r, 512, 512)
        gemm[name] = {"M": M, "N": N, "K": K, "product": kind, "tile": tile, "ms": ms,
                      "tflops": 2.0 * M * N * K / (ms / 1e3) / 1e12}
    gemm["hipblaslt_65536_rows"] = hipblaslt_gemm(65536, 512, 512)
    rng = np.random.default_rng(0)
    x = rng.uniform(-1, 1, size=(4096, surrogate.N_FEATURES)).astype(np.float32)
    y = rng.uniform(-1, 1, size=4096).astype(np.float32)
    trainer.train_step(x, y, 1e-4)
    t1 = time.perf_counter()
    for _ in range(50):
        trainer.train_step(x, y, 1e-4)
    step_ms = (time.perf_counter() - t1) * 1e3 / 50
    trainer.close()
    return gemm, step_ms


def gemm_roofline(gemm: dict) -> dict:
    dom = gemm["forward_hidden_65536_rows"]
    # the kernel the autotuner picked; traffic only for the kernels whose counters were measured
    traffic, kernel, source = GEMM_FWD_65536_TRAFFIC.get(
        dom["tile"], (None, "gemm_nt_bres_kernel" if "resident" in dom["tile"] else "gemm_nt_kernel", None))
    return {"bound": "mfma", "kernel": kernel,
            "shape": "65536 x 512 x 512 (forward of a hidden layer, bias + ReLU + bf16 out fused)",
            "achieved": dom["tflops"], "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": dom["tflops"] / MFMA_BF16_PEAK_TFLOPS, "tile": dom["tile"],
            # HBM bytes per launch (FETCH_SIZE + WRITE_SIZE scaled by the copy-kernel calibration)
            "traffic": traffic,
            "traffic_copied_from": f"{source} (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE)" if source else None,
            "algorithmic_bytes": 2.0 * (65536 * 512 + 512 * 512 + 65536 * 512)}


def gemm_bench(args) -> dict:
    """The surrogate's MFMA products alone (no Monte Carlo): what the gemm_nt_kernel and
    gemm_nt_bres_kernel profiles measure."""
    import torch  # noqa: F401  (before the library initialises HIP)

    t0 = time.perf_counter()
    gemm, step_ms = surrogate_products()
    dom = gemm["forward_hidden_65536_rows"]
    return {"metric": "surrogate forward product TFLOP/s", "value": dom["tflops"], "unit": "TFLOP/s", "n_gpus": 1,
            "steps": 1, "warmup": 0, "ms_per_step": (time.perf_counter() - t0) * 1e3, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "bf16 (fp32 accumulate)", "data": "synthetic",
            "config": {"workload": "surrogate 32 -> 512 x 4 -> 1: fused products of a hidden layer at 4096 and "
                                   "65536 rows, and a training step at batch 4096", "parallelism": "single GPU"},
            "roofline": gemm_roofline(gemm), "gemm": gemm, "training_step_ms_b4096": step_ms, "cpu_baseline": None}


def surrogate_bench(args) -> dict:
    """Warm start of held-out structures from the trained surrogate against the cold start, and the
    surrogate's MFMA GEMM against the bf16 peak."""
    import numpy as np
    import torch  # noqa: F401  (before the engine initialises HIP: torch's own device setup then succeeds)
    import psim_amd
    from psim_amd import configs, surrogate, train_surrogate as ts

    count = args.structures or 128
    params = ts.structures(ts.EVAL_SEED, count)
    engine = psim_amd.Engine(0)
    net = surrogate.load_weights()
    phonons, steps, sim_ns = 8_000_000, 5000, 25.0
    t0 = time.perf_counter()
    res = ts.evaluate(engine, net, params, phonons, steps, sim_ns, seed=7, linear_count=min(32, count),
                      progress=lambda m: print(m, file=sys.stderr, flush=True))
    wall = time.perf_counter() - t0
    s = res["summary"]
    row = s["criteria"][str(ts.HEADLINE_TOL)]
    rows = sum(p["n_sensors"] for p in params)
    gemm, step_ms = surrogate_products(rows)
    line = {
        "metric": "time-to-steady-state cold/warm", "value": row["cold_over_warm_total"], "unit": "x",
        "n_gpus": 1, "steps": 1, "warmup": 0, "ms_per_step": wall * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "bf16 network (fp32 accumulate), f64 Monte Carlo",
        "data": "synthetic",
        "config": {"workload": f"{count} held-out structures (psim_amd.train_surrogate.structures, seed "
                               f"{ts.EVAL_SEED}), each from a cold start and from the surrogate's prediction, "
                               f"{phonons} phonons, {steps} steps of {sim_ns / steps * 1e3:g} ps, periodic mode",
                   "criterion": f"RMS over sensors of the change in 100-step window-mean T < {ts.HEADLINE_TOL} x "
                                f"|t_hot - t_cold| for 3 windows",
                   "network": "32 -> 512 x 4 (ReLU) -> 1, psim_amd/data/surrogate_32x512x4.npz",
                   "parallelism": "single GPU"},
        "cold_over_warm_gpu_seconds": s["gpu_seconds_to_steady_state"]["cold_over_warm"],
        "time_to_steady_state": s,
        "roofline": gemm_roofline(gemm),
        "gemm": gemm,
        "training_step_ms_b4096": step_ms,
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        sample = [configs.geometry_from_params(dict(p, phonons=20_000), steps, sim_ns, sim_type=0) for p in params[:2]]
        rates = [reference_rate(c) for c in sample]
        counted = sum(r[2] for r in rates)
        wall_cpu = sum(r[1] for r in rates)
        line["cpu_baseline"] = {"value": counted / wall_cpu, "unit": "phonon-steps/s", "cores": 1,
                                "kind": rates[0][3],
                                "sample": "the first 2 held-out structures, cold start, 2e4 phonons each, "
                                          f"{steps} steps of {sim_ns / steps * 1e3:g} ps, steady-state mode, one "
                                          "process; phonon-steps counted by the oracle"}
    engine.close()
    return line


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", choices=["lshape", "film", "batch", "surrogate", "gemm"], default="lshape")
    ap.add_argument("--structures", type=int, default=None, help="held-out structures (surrogate workload)")
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--phonons", type=int, default=None)
    ap.add_argument("--cpu-sample-phonons", type=int, default=None)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-steady-state", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="batch workload: keep the default launch shape (for kernel traces of batched runs only)")
    ap.add_argument("--no-sequential", action="store_true",
                    help="batch workload: skip the one-by-one runs of the same models")
    args = ap.parse_args()
    film = args.workload == "film"
    if args.steps is None:
        args.steps = 1 if film else 3
    if args.warmup is None:
        args.warmup = 0 if film else 1
    if args.phonons is None:
        args.phonons = 100_000_000 if film else 10_000_000
    if args.cpu_sample_phonons is None:
        args.cpu_sample_phonons = 10_000 if film else 100_000
    if film:
        os.environ.setdefault("PSIM_PROGRESS", "1")  # one stderr line per launch window
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: this build runs on a single MI355X (the multi-GPU design, sharding "
              "phonon ids and all-reducing the tallies, is not run here).",
              file=sys.stderr)
        return 2

    if args.workload == "batch":
        print(json.dumps(batch_bench(args)), flush=True)
        return 0
    if args.workload == "surrogate":
        print(json.dumps(surrogate_bench(args)), flush=True)
        return 0
    if args.workload == "gemm":
        print(json.dumps(gemm_bench(args)), flush=True)
        return 0

    import psim_amd

    cfg, desc = workload(args.workload, args.phonons)
    engine = psim_amd.Engine(0)
    model = psim_amd.Model(cfg)
    _, total = model.upload(engine, 0)
    for w in range(args.warmup):
        engine.run_resident(10_000 + w)
    steps_total = 0
    kernel_ms = 0.0
    launches = 0
    t0 = time.perf_counter()
    flux64 = None
    for k in range(args.steps):
        engine.run_resident(k)
        _, flux, counters = engine.fetch(model.num_sensors, model.num_record)
        if k == 0:
            flux64 = flux
        tm = engine.last_timing()
        steps_total += counters["phonon_steps"]
        kernel_ms += tm["step_kernel_ms"]
        launches += tm["step_launches"]
    wall = time.perf_counter() - t0
    value = steps_total / wall
    achieved = BYTES_PER_PHONON_STEP * steps_total / (kernel_ms / 1e3) / 1e9
    line = {
        "metric": "phonon-steps/s",
        "value": value,
        "unit": "phonon-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": dict(desc, phonons=total),
        "method": "phonons fly across the measurement boundaries; each boundary passed is counted as a phonon-step "
                  "and, when recorded, tallied with the phonon's sensor and velocity of that moment",
        "roofline": {
            # the step-synchronous HBM model of this path; what the counters show limits the
            # kernel (VALU issue) is `limiter` below
            "bound": "hbm",
            "model": "120 algorithmic bytes per phonon-step: one 60-byte SoA phonon record read and written "
                     "per measurement step",
            "kernel": "transport_kernel",
            "achieved": achieved,
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBPS,
            "traffic": None,
            "bytes_per_phonon_step": BYTES_PER_PHONON_STEP,
            "kernel_ms_per_launch": kernel_ms / max(1, launches),
            "launches": launches,
        },
        "cpu_baseline": None,
    }
    # tallies: boundaries recorded per run and per second of transport time; a flight records the
    # boundaries it passes as one difference-array segment (six atomics whatever its length)
    line["tallies"] = {"boundaries_per_run": counters["tallies"],
                       "boundaries_per_s_of_transport_time": counters["tallies"] / (kernel_ms / 1e3 / args.steps)}
    # What the counters say of this kernel on this workload (rocprofv3 --pmc, profiles/): HBM bytes
    # per phonon-step (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration) and VALU
    # issue.  Counters cannot be read inside this process, so these figures are copied from the
    # profile, labelled with the engine they were measured on and whether it is this one.
    if not film and args.phonons == 10_000_000:
        copied = profiled_counters()
        if copied is not None:
            per_launch = steps_total / max(1, launches)
            line["roofline"]["traffic"] = copied["traffic"]["bytes_per_phonon_step"] * per_launch
            line["roofline"]["traffic_bytes_per_phonon_step"] = copied["traffic"]["bytes_per_phonon_step"]
            line["roofline"]["limiter"] = copied["limiter"]
            line["roofline"]["issue"] = copied["issue"]
            line["roofline"]["counters_copied_from"] = {
                "file": "profiles/transport_kernel_lshape_1e7_counters.json",
                "engine_commit": copied["engine_commit"], "engine_is_current": copied["current"],
                "traffic": "bytes per phonon-step from the profile x this run's phonon-steps per launch"}
    # The same run (seed 0) with FP32 flux tallies summed with Kahan compensation, against FP64:
    # the same trajectories, so the difference is the FP32 tallies' error alone.
    if flux64 is not None:
        engine.set_tally_precision("fp32_kahan")
        engine.run_resident(0)
        _, flux32, c32 = engine.fetch(model.num_sensors, model.num_record)
        ms32 = engine.last_timing()["run_ms"]
        engine.set_tally_precision("fp64")
        m32, m64 = flux32.mean(axis=1), flux64.mean(axis=1)  # [sensor][x, y] mean over the recorded steps
        line["fp32_kahan_tallies"] = {
            "value": c32["phonon_steps"] / (ms32 / 1e3), "unit": "phonon-steps/s", "run_ms": ms32,
            "max_abs_diff_over_max_flux": float(abs(flux32 - flux64).max() / max(abs(flux64).max(), 1e-300)),
            "max_diff_sensor_mean_flux_over_largest": float(abs(m32 - m64).max() / max(abs(m64).max(), 1e-300)),
            "method": "FP32 atomics into the LDS/HBM difference arrays, Kahan-compensated FP32 sums along the steps "
                      "and across blocks, returned as sum - compensation; compared with the FP64 run of the same seed"}
    # the boundary's own call end to end: host tables and budget, upload, run, tally copy-back and
    # interpretation (psim_model_run); reported beside `value`, which has the model resident in HBM
    if not film:
        t0 = time.perf_counter()
        model.run(engine, 12_345)
        e2e = time.perf_counter() - t0
        e2e_steps = model.results(0).counters["phonon_steps"]
        line["end_to_end"] = {"wall_s": e2e, "phonon_steps": e2e_steps, "rate": e2e_steps / e2e,
                              "call": "psim_model_run"}
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(args.workload, args.cpu_sample_phonons)
    if not args.no_steady_state and not film:
        line["time_to_steady_state"] = steady_state(engine, 10_000_000, not args.no_cpu_baseline)
    print(json.dumps(line), flush=True)
    engine.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
