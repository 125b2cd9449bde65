#!/usr/bin/env python3
"""Throughput of the island-model filter step on one MI355X.

Workload (plain form): unicycle model, 32 range+bearing landmarks drawn
uniformly in [-50, 50]^2 from seed 42, 4096 sub-filters x 512 particles
= 2,097,152 particles (FP32 SoA state), 64x64 torus exchanging the 4 best
particles per neighbour, MTGP32 seed 42 with the reference's 96-stream buffer,
log-domain weights, resampling every step.  Two resamplers: prefix-sum
systematic (the reported value) and Vose alias.

One step = MTGP32 words -> sample+weight (Box-Muller fused) -> bitonic sort +
gather -> argmax + posterior -> torus exchange -> resampling, all on the GPU;
inputs (particles, RNG state) are resident in HBM; the per-step measurement
frame (~1.3 KB) is copied host->device inside the step.

Prints one JSON line (see the keys below).  --gpus N with N > 1 exits: this
build runs on one GPU.
"""
import argparse
import ctypes
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

METRIC = "particles/s per step @2M unicycle (% HBM roofline); batched p50/p99 step latency"
HBM_PEAK_GBS = 8000.0
M, G, T, L = 512, 4096, 4, 32
S = 3
# the dominant launch, sample+weight and block sort fused (k_sample_sort):
# algorithmic bytes per particle: state read 12 + uniforms read 12 + sorted
# log-weight write 4 + sorted state write 12; per sub-filter: the gate's
# uniform 4, winner 8, posterior partials 4 (2 + S), compact best t 4 t (1 + S)
SS_BYTES_PER_PARTICLE = 40
SS_BYTES_PER_SUB = 4 + 8 + 4 * (2 + S) + 4 * T * (1 + S)
SS_BYTES_PER_LAUNCH = SS_BYTES_PER_PARTICLE * M * G + SS_BYTES_PER_SUB * G
# k_sample_weight alone: state read 12 + uniforms read 12 + state write 12 +
# log-weight write 4 bytes per particle
SW_BYTES_PER_LAUNCH = 40 * M * G
STAGES = ["rng_wait", "sample_weight", "sort", "argmax", "exchange_or_topt", "resample"]


def stage_dict(us, fused):
    """Per-stage device times; with a scan resampler the argmax runs inside the
    resampling launch, so the two are reported together."""
    d = {k: v for k, v in zip(STAGES, us) if v == v}
    if fused:
        d.pop("argmax", None)
        if "resample" in d:
            d["resample_with_argmax"] = d.pop("resample")
    return d


def settle(ea, eng, seconds):
    """Untimed steps of the headline filter (a filter of its own) for about
    `seconds`: the GPU raises its engine clock over the first tens of
    milliseconds of load, and a measurement of a few steps taken before that
    measures the clock ramp (0.129 against 0.113 ms per step at 20 steps)."""
    if seconds <= 0:
        return
    import numpy as np
    desc = ea.filter_desc(ea.MODEL_UNICYCLE, m=M, G=G, t=T, topology=ea.TOPO_TORUS, resampler=ea.RESAMPLE_SYSTEMATIC,
                          weights=ea.WEIGHTS_LOG, rng=ea.RNG_MTGP_COMPAT, seed=42)
    params, ctrl, meas, truth = ea.unicycle_scenario(42, L, 50.0, 64)
    f = ea.Filter(eng, desc, params)
    f.init_uniform(np.zeros(3, np.float32))
    t0 = time.perf_counter()
    k = 0
    while time.perf_counter() - t0 < seconds:
        for _ in range(32):
            f.step(ctrl[k % 64], meas[k % 64], 0.1)
            k += 1
        f.sync()
    f.close()


def run_plain(ea, eng, resampler, steps, warmup, graph=False, timing=1, every=1, landmarks=L, box=50.0,
              fusion=True):
    """One pass of the headline filter: `warmup` untimed steps, then `steps`
    timed ones.  timing 0: no events in the stream (the throughput pass);
    1: events around the sample+weight(+sort) launch of one step in `every`;
    2: events around every stage."""
    desc = ea.filter_desc(ea.MODEL_UNICYCLE, m=M, G=G, t=T, topology=ea.TOPO_TORUS, resampler=resampler,
                          weights=ea.WEIGHTS_LOG, rng=ea.RNG_MTGP_COMPAT, seed=42)
    params, ctrl, meas, truth = ea.unicycle_scenario(42, landmarks, box, warmup + steps)
    f = ea.Filter(eng, desc, params)
    import numpy as np
    f.init_uniform(np.zeros(3, np.float32))
    if f.set_fusion(fusion) != fusion:
        raise RuntimeError("the fused sample+sort launch is expected at m = 512 for the unicycle")
    if graph:
        f.use_graph(True)
    for k in range(warmup):
        f.step(ctrl[k], meas[k], 0.1)
    ea.device_sync()
    if not graph and timing:
        f.set_timing(timing, every)
    t0 = time.perf_counter()
    for k in range(warmup, warmup + steps):
        f.step(ctrl[k], meas[k], 0.1)
    ea.device_sync()
    t1 = time.perf_counter()
    stage = [float(x) for x in f.stage_times()] if (timing and not graph) else [float("nan")] * 6
    est = f.estimate()
    k = warmup + steps - 1
    err = float(np.hypot(est.mean[0] - truth[k, 0], est.mean[1] - truth[k, 1]))
    f.close()
    return t1 - t0, stage, err


def run_layouts(ea, eng):
    """Device time per step of the headline filter (unicycle, S = 3) and of the
    robot arm at the same size (S = 9) with the state resident as SOA, AOS
    and PACKED (the layout autotuner's measurement)."""
    import numpy as np
    out = {}
    names = ["soa", "aos", "packed"]
    params, ctrl, meas, truth = ea.unicycle_scenario(42, L, 50.0, 1)
    d = ea.filter_desc(ea.MODEL_UNICYCLE, m=M, G=G, t=T, topology=ea.TOPO_TORUS, resampler=ea.RESAMPLE_SYSTEMATIC,
                       weights=ea.WEIGHTS_LOG, rng=ea.RNG_MTGP_COMPAT, seed=42)
    us, best = ea.autotune_layout(eng, d, params, ctrl[0], meas[0], 0.1)
    out["unicycle_2m"] = {"us_per_step": dict(zip(names, map(float, us))), "best": names[best]}
    c = ea.RobotArm5Control()
    m = ea.RobotArm5Measurement()
    for i in range(5):
        c.angles[i] = 0.05
        m.angles[i] = 2.35
    m.x, m.y = 1.0, 2.0
    d = ea.filter_desc(ea.MODEL_ROBOTARM5, m=M, G=G, t=T, topology=ea.TOPO_TORUS, resampler=ea.RESAMPLE_SYSTEMATIC,
                       weights=ea.WEIGHTS_LOG, rng=ea.RNG_MTGP_COMPAT, seed=42)
    us, best = ea.autotune_layout(eng, d, None, c, m, 0.0589)
    out["robotarm_2m"] = {"us_per_step": dict(zip(names, map(float, us))), "best": names[best]}
    return out


def run_accuracy(ea, eng, steps=500):
    """Position RMSE against the seeded ground truth: the small unicycle
    workload (16 x 256 particles, 4 landmarks at (+-10, +-10), 4 x 4 torus,
    t = 1, systematic resampling, seed 1234) over 500 steps, posterior mean
    and best particle."""
    import numpy as np
    d = ea.filter_desc(ea.MODEL_UNICYCLE, m=256, G=16, t=1, topology=ea.TOPO_TORUS,
                       resampler=ea.RESAMPLE_SYSTEMATIC, weights=ea.WEIGHTS_LOG, seed=1234)
    p, ctrl, meas, truth = ea.unicycle_scenario(1234, 4, 0.0, steps)
    f = ea.Filter(eng, d, p)
    f.init_uniform(np.zeros(3, np.float32))
    em, eb = [], []
    for k in range(steps):
        f.step(ctrl[k], meas[k], 0.1)
        est = f.estimate()
        em.append(np.hypot(est.mean[0] - truth[k, 0], est.mean[1] - truth[k, 1]))
        eb.append(np.hypot(est.best[0] - truth[k, 0], est.best[1] - truth[k, 1]))
    f.close()
    return {"workload": "unicycle 16 x 256, 4 landmarks, 4x4 torus t=1, systematic, seed 1234, 500 steps",
            "rmse_posterior_mean": float(np.sqrt(np.mean(np.square(em)))),
            "rmse_best_particle": float(np.sqrt(np.mean(np.square(eb)))),
            "max_error_posterior_mean": float(np.max(em))}


def run_accuracy_headline(ea, eng, steps=1000):
    """The 2M-particle headline filter over its 1000 steps against the seeded
    ground truth: RMSE of the posterior mean (position, heading) and of the
    best particle's position."""
    import numpy as np
    d = ea.filter_desc(ea.MODEL_UNICYCLE, m=M, G=G, t=T, topology=ea.TOPO_TORUS, resampler=ea.RESAMPLE_SYSTEMATIC,
                       weights=ea.WEIGHTS_LOG, rng=ea.RNG_MTGP_COMPAT, seed=42)
    p, ctrl, meas, truth = ea.unicycle_scenario(42, L, 50.0, steps)
    f = ea.Filter(eng, d, p)
    f.init_uniform(np.zeros(3, np.float32))
    ep, eh, eb = [], [], []
    for k in range(steps):
        f.step(ctrl[k], meas[k], 0.1)
        e = f.estimate()
        ep.append(np.hypot(e.mean[0] - truth[k, 0], e.mean[1] - truth[k, 1]))
        eh.append(np.angle(np.exp(1j * (e.mean[2] - truth[k, 2]))))
        eb.append(np.hypot(e.best[0] - truth[k, 0], e.best[1] - truth[k, 1]))
    f.close()
    return {"workload": f"unicycle {G} x {M}, {L} landmarks, 64x64 torus t={T}, systematic, seed 42, {steps} steps",
            "rmse_position_posterior_mean": float(np.sqrt(np.mean(np.square(ep)))),
            "rmse_heading_posterior_mean": float(np.sqrt(np.mean(np.square(eh)))),
            "rmse_position_best_particle": float(np.sqrt(np.mean(np.square(eb)))),
            "max_position_error_posterior_mean": float(np.max(ep))}


DH_M, DH_G, DH_T = 512, 2048, 1
# k_sample_weight<dh6>: state read 48 + uniforms read 24 (6 per particle) +
# state write 48 + log-weight write 4 bytes per particle
DH_BYTES_PER_PARTICLE = 124


def run_dh6(ea, eng, steps, warmup, acc_steps=2000):
    """The robot-arm configuration: 6-joint DH arm (seed 7), 2048 sub-filters x
    512 = 1,048,576 particles on the 32 x 64 torus, t = 1, systematic
    resampling, dt = 0.01.  Throughput (untimed stream), the sample+weight
    kernel's time from events around it on every step of a separate pass, and
    the RMSE of the posterior mean against the seeded truth over 2000 steps."""
    import numpy as np
    d = ea.filter_desc(ea.MODEL_DH6, m=DH_M, G=DH_G, t=DH_T, topology=ea.TOPO_TORUS,
                       resampler=ea.RESAMPLE_SYSTEMATIC, weights=ea.WEIGHTS_LOG, seed=7)
    ksteps = min(max(steps, 50), 200)
    n = max(steps, ksteps, acc_steps) + warmup
    p, meas, truth = ea.dh6_scenario(7, n)
    c = ea.DH6Control()
    N = DH_M * DH_G

    def one(timed_steps, timing):
        f = ea.Filter(eng, d, p)
        f.init_uniform(truth[0])
        for k in range(warmup):
            f.step(c, meas[k], 0.01)
        f.sync()
        if timing:
            f.set_timing(1, 1)
        t0 = time.perf_counter()
        for k in range(warmup, warmup + timed_steps):
            f.step(c, meas[k], 0.01)
        f.sync()
        secs = time.perf_counter() - t0
        us = float(f.stage_times()[1]) if timing else None
        f.close()
        return secs, us

    secs, _ = one(steps, False)
    _, kus = one(ksteps, True)
    f = ea.Filter(eng, d, p)
    f.init_uniform(truth[0])
    eq, ev = [], []
    for k in range(acc_steps):
        f.step(c, meas[k], 0.01)
        e = f.estimate()
        mean = np.array(e.mean[:12])
        eq.append(np.sqrt(np.mean(np.square(mean[:6] - truth[k + 1, :6]))))
        ev.append(np.sqrt(np.mean(np.square(mean[6:] - truth[k + 1, 6:]))))
    f.close()
    a = DH_BYTES_PER_PARTICLE * N / (kus * 1e-6) / 1e9
    return {"workload": f"dh6 arm (seed 7), {DH_G} sub-filters x {DH_M} = {N} particles, 32x64 torus t={DH_T}, "
                        f"systematic resampling, dt 0.01",
            "particles": N, "steps": steps, "ms_per_step": secs * 1e3 / steps, "particles_per_s": N * steps / secs,
            "roofline": {"bound": "hbm", "kernel": "k_sample_weight<dh6, SOA>", "kernel_us": kus,
                         "bytes_per_launch": DH_BYTES_PER_PARTICLE * N, "achieved": a, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": a / HBM_PEAK_GBS},
            "accuracy": {"steps": acc_steps,
                         "rmse_joint_angle_rad": float(np.sqrt(np.mean(np.square(eq)))),
                         "rmse_joint_velocity_rad_s": float(np.sqrt(np.mean(np.square(ev))))}}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(warm=3, reps=5, rep_s=3.0):
    """The oracle's FP32 restatement of the reference step (the same MTGP
    stream, bitonic network, scan and resampler), OpenMP over sub-filters,
    built here with -O3 -march=native (make -C oracle baseline), on a bounded
    sample of the workload: the full 4096 x 512 filter, 3 untimed steps, then
    5 repetitions of the step count that takes about `rep_s` seconds; the
    median repetition is reported."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import statistics
    import subprocess
    import tempfile

    import numpy as np

    import esthera_amd as ea
    import oracle
    out = tempfile.mkdtemp(prefix="esth_cpu_baseline_")
    r = subprocess.run(["make", "-s", "-C", os.path.join(REPO, "oracle"), "baseline", f"OUT={out}"],
                       capture_output=True, text=True)
    native = os.path.join(out, "liboracle_native.so")
    build = "gcc -O3 -march=native -fopenmp (oracle/Makefile baseline)"
    if r.returncode != 0 or not os.path.exists(native):
        native, build = None, "the committed -O2 oracle build (the native build failed)"
    olib = oracle.use_library(native) if native else oracle.lib()
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    olib.or_set_threads(threads)
    desc = ea.filter_desc(ea.MODEL_UNICYCLE, m=M, G=G, t=T, topology=ea.TOPO_TORUS,
                          resampler=ea.RESAMPLE_SYSTEMATIC, weights=ea.WEIGHTS_LOG, rng=ea.RNG_MTGP_COMPAT, seed=42)
    nrow = 400
    params, ctrl, meas, truth = ea.unicycle_scenario(42, L, 50.0, nrow)
    o = oracle.Filter(desc, params, ctypes.sizeof(params))
    o.set_f64(False)
    o.init_uniform(np.zeros(3, np.float32))
    k = 0
    t0 = time.perf_counter()
    for _ in range(warm):
        o.step(ctrl[k % nrow], meas[k % nrow], 0.1)
        k += 1
    per_step = (time.perf_counter() - t0) / warm
    n = max(1, int(rep_s / per_step))
    rates = []
    for _ in range(reps):
        t0 = time.perf_counter()
        for _ in range(n):
            o.step(ctrl[k % nrow], meas[k % nrow], 0.1)
            k += 1
        rates.append(M * G * n / (time.perf_counter() - t0))
    return {"value": statistics.median(rates), "unit": "particles/s", "cores": threads, "kind": "port",
            "cpu": cpu_model(), "nproc": os.cpu_count(), "build": build,
            "sample": f"full {G} x {M} workload, {warm} untimed steps, then the median of {reps} repetitions of "
                      f"{n} steps (rates {min(rates):.3g}..{max(rates):.3g} particles/s), FP32 oracle "
                      f"restatement, OpenMP {threads} threads"}


def run_batched(ea, eng, steps=200, warmup=3, nf=10000, plain_every=20):
    """10,000 unicycle filters with 8 landmarks each (64 seeded landmark sets),
    particle counts log-uniform in [256, 65536] rounded to multiples of 64
    (seed 99), ring exchange t=1, systematic resampling, sub-filter size per
    filter from the autotuner; the whole batch stepped by one graph launch per
    step.  Beside it, every `plain_every`-th filter stepped as its own filter
    (per-filter launches), and the two compared filter by filter."""
    import numpy as np

    from esthera_amd import workloads
    scen = workloads.batched_scenarios()
    NS = len(scen)
    p0, c0, m0, _ = scen[0]
    d0 = ea.filter_desc(ea.MODEL_UNICYCLE, m=256, G=4, t=1, topology=ea.TOPO_RING, resampler=ea.RESAMPLE_SYSTEMATIC,
                        weights=ea.WEIGHTS_LOG, rng=ea.RNG_MTGP_PER_FILTER, seed=1)
    us = ea.autotune(eng, d0, p0, c0[0], m0[0], 0.1)
    counts, descs, params, scen = workloads.batched_filters(us, nf=nf, scen=scen)
    b = ea.Batch(eng, descs, params)
    b.init_uniform(np.zeros((nf, 3), np.float32))
    IN = [workloads.batched_inputs(scen, nf, k) for k in range(workloads.BATCH_SCENARIO_STEPS)]
    dts = np.full(nf, 0.1, np.float32)
    for k in range(warmup):
        b.step(IN[k % 8][0], IN[k % 8][1], dts)
    b.sync()
    b.step_times()
    t0 = time.perf_counter()
    for k in range(warmup, warmup + steps):
        b.step(IN[k % 8][0], IN[k % 8][1], dts)
    b.sync()
    secs = time.perf_counter() - t0
    st = b.step_times() / 1e3
    sub = list(range(0, nf, plain_every))
    fs = []
    for i in sub:
        f = ea.Filter(eng, descs[i], params[i])
        f.init_uniform(np.zeros(3, np.float32))
        fs.append(f)
    for k in range(warmup):
        for j, f in enumerate(fs):
            f.step(scen[sub[j] % NS][1][k % 8], scen[sub[j] % NS][2][k % 8], 0.1)
    ea.device_sync()
    t0 = time.perf_counter()
    for k in range(warmup, warmup + steps):
        for j, f in enumerate(fs):
            f.step(scen[sub[j] % NS][1][k % 8], scen[sub[j] % NS][2][k % 8], 0.1)
    ea.device_sync()
    psecs = time.perf_counter() - t0
    worst, same = 0.0, 0
    for j, f in enumerate(fs):
        a, e = b.estimate(sub[j]), f.estimate()
        ma, me = np.array(a.mean[:3]), np.array(e.mean[:3])
        worst = max(worst, float(np.max(np.abs(ma - me) / np.maximum(np.abs(me), 1.0))))
        sa, _ = b.particles(sub[j])
        se, _ = f.particles()
        same += int(np.array_equal(sa.view(np.uint32), se.view(np.uint32)))
        f.close()
    total = int(counts.sum())
    ms = [d.num_particles for d in descs]
    b.close()
    plain_particles = int(sum(int(counts[i]) for i in sub))
    return {
        "workload": f"{nf} unicycle filters, 8 landmarks each, particles log-uniform in [256, 65536] (x64, "
                    f"seed 99), ring t=1, systematic resampling, per-filter MTGP32 streams, one HIP graph per step",
        "filters": nf, "particles": total, "steps": steps,
        "sub_filter_sizes": {str(m): ms.count(m) for m in sorted(set(ms))},
        "autotune_us_per_mparticle": {str(64 << k): float(us[k]) for k in range(5)},
        "ms_per_step": secs * 1e3 / steps,
        "particles_per_s": total * steps / secs,
        "filter_updates_per_s": nf * steps / secs,
        "step_latency_ms": {"p50": float(np.percentile(st, 50)), "p99": float(np.percentile(st, 99)),
                            "max": float(np.max(st))},
        "plain": {"filters": len(sub), "particles": plain_particles, "ms_per_round": psecs * 1e3 / steps,
                  "filter_updates_per_s": len(sub) * steps / psecs,
                  "particles_per_s": plain_particles * steps / psecs},
        "speedup_filter_updates": (nf * steps / secs) / (len(sub) * steps / psecs),
        "parity_vs_plain": {"filters_compared": len(sub), "bit_identical_particles": same,
                            "max_rel_mean_diff": worst},
    }


def run_served(ea, eng, rates=(1000.0, 5000.0, 20000.0, 50000.0), duration=2.0, tracks=512):
    """Poisson request streams (5 % opens: 70 % unicycle with 8 landmarks / 30 %
    robot arm replaying input5, particles log-uniform in [1024, 262144]; 95 %
    updates of a uniformly chosen open track; geometric lifetime of mean 500
    updates; HBM arena 200 GB or what the device holds) at each rate, then the
    closed-loop capacity of the server against plain per-request launches on
    the same `tracks` tracks."""
    import numpy as np
    meas, ctrl, dts, _ = ea.load_robotarm_trace(os.path.join(REPO, "tests", "golden", "robotarm_input5.txt"))
    p, c, m, _ = ea.unicycle_scenario(1000, 8, 20.0, 2)
    d0 = ea.filter_desc(ea.MODEL_UNICYCLE, m=256, G=4, t=1, topology=ea.TOPO_RING, resampler=ea.RESAMPLE_SYSTEMATIC,
                        weights=ea.WEIGHTS_LOG, rng=ea.RNG_MTGP_PER_FILTER, seed=1)
    us = ea.autotune(eng, d0, p, c[0], m[0], 0.1)
    used, budget = eng.memory()
    arena = int(min(200 * 10 ** 9, 0.7 * (budget - used)))

    def workload(rate, initial, lifetime, plain, dur):
        return ea.ServeWorkload(rate=rate, duration_s=dur, open_fraction=0.05, unicycle_fraction=0.7,
                                min_particles=1024, max_particles=262144, mean_lifetime=lifetime,
                                initial_tracks=initial, landmarks=8, topology=ea.TOPO_TORUS, num_transfer=1,
                                resampler=ea.RESAMPLE_SYSTEMATIC, weights=ea.WEIGHTS_LOG, seed=4, plain=plain,
                                hbm_bytes=arena, max_tracks=1 << 20, max_inflight=2)

    keys = ("updates_per_s", "p50_us", "p99_us", "max_us", "mean_wave_updates", "peak_tracks", "hbm_peak",
            "backlog", "failed", "mean_particles")
    sweep = []
    for rate in rates:
        r = ea.serve_simulate(eng, workload(rate, 256, 500.0, 0, duration), (meas, ctrl, dts), us)
        sweep.append(dict({"rate": rate}, **{k: r[k] for k in keys}))
    cap = ea.serve_simulate(eng, workload(0.0, tracks, 0.0, 0, duration), (meas, ctrl, dts), us)
    plain = ea.serve_simulate(eng, workload(0.0, tracks, 0.0, 1, duration), (meas, ctrl, dts), us)
    return {
        "workload": "Poisson requests: 5% open (70% unicycle 8 landmarks / 30% robot arm on input5; particles "
                    "log-uniform [1024, 262144]), 95% update, geometric lifetime mean 500 updates",
        "hbm_arena_bytes": arena,
        "duration_s": duration,
        "sweep": sweep,
        "capacity": {"tracks": tracks,
                     "served": {k: cap[k] for k in ("updates_per_s", "particle_updates_per_s", "p50_us", "p99_us",
                                                    "mean_wave_updates")},
                     "plain": {k: plain[k] for k in ("updates_per_s", "particle_updates_per_s", "p50_us",
                                                     "p99_us")},
                     "speedup_updates": cap["updates_per_s"] / max(plain["updates_per_s"], 1e-9)},
    }


def traffic_from_profiles():
    p = os.path.join(REPO, "profiles", "sample_sort_unicycle2m_traffic.json")
    if not os.path.exists(p):
        return None
    with open(p) as fh:
        return json.load(fh).get("bytes_per_launch")


# VALU issue ceiling of MI355X: 1024 SIMDs (256 CUs x 4), each issuing at most
# two VALU wave-instructions per quad-cycle, at the 2.4 GHz peak engine clock
VALU_PEAK_WAVE_INST_S = 1024 * 2.4e9 / 2


def valu_from_profiles(kernel_us):
    """VALU issue rate of k_sample_sort: the wave-instructions per launch
    counted by rocprofv3 (profiles/sample_sort_unicycle2m_valu.md) over the
    launch time measured here, against the dual-issue ceiling."""
    p = os.path.join(REPO, "profiles", "sample_sort_unicycle2m_valu.json")
    if not os.path.exists(p) or not kernel_us:
        return None
    with open(p) as fh:
        ins = json.load(fh)["valu_wave_insts_per_launch"]
    achieved = ins / (kernel_us * 1e-6)
    return {"bound": "valu", "achieved": achieved / 1e12, "peak": VALU_PEAK_WAVE_INST_S / 1e12,
            "unit": "T wave-instructions/s", "frac": achieved / VALU_PEAK_WAVE_INST_S,
            "wave_insts_per_launch": ins}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-alias", action="store_true", help="skip the Vose alias variant")
    ap.add_argument("--graph-only", action="store_true", help="only the HIP-graph replay run (profiling)")
    ap.add_argument("--main-only", action="store_true",
                    help="only the timed headline pass and its roofline (profiling: one workload per kernel)")
    ap.add_argument("--no-batched", action="store_true", help="skip the 10k-filter batched workload")
    ap.add_argument("--no-served", action="store_true", help="skip the served request-stream workload")
    ap.add_argument("--no-layouts", action="store_true", help="skip the state-layout comparison")
    ap.add_argument("--no-dh6", action="store_true", help="skip the robot-arm (dh6, 2048 x 512) workload")
    ap.add_argument("--no-accuracy", action="store_true", help="skip the accuracy passes")
    ap.add_argument("--settle", type=float, default=0.3,
                    help="seconds of untimed headline steps before the warmup (engine clock ramp)")
    ap.add_argument("--copy-calib", type=int, default=0,
                    help="also run one device copy of this many MiB (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              f"(sharding across GPUs is not part of it)", file=sys.stderr)
        sys.exit(2)

    import esthera_amd as ea
    eng = ea.Engine(0)
    if args.copy_calib:
        nbytes = args.copy_calib << 20
        a, b = ea.DeviceBuffer(nbytes), ea.DeviceBuffer(nbytes)
        ea.check(ea.lib().esth_memset_d(a.ptr, 1, nbytes))
        ea.check(ea.lib().esth_stage_copy(b.ptr, a.ptr, nbytes))
        a.free()
        b.free()

    N = M * G
    if args.graph_only:
        gsecs, _, _ = run_plain(ea, eng, ea.RESAMPLE_SYSTEMATIC, args.steps, args.warmup, graph=True)
        print(json.dumps({"graph_ms_per_step": gsecs * 1e3 / args.steps}), flush=True)
        return
    settle(ea, eng, args.settle)
    # timed region: eager steps, no events in the stream
    secs, _, err = run_plain(ea, eng, ea.RESAMPLE_SYSTEMATIC, args.steps, args.warmup, timing=0)
    value = N * args.steps / secs
    # the dominant launch's own time: a separate pass of 400 steps with events
    # around the sample+weight+sort launch of every 10th step (40 launches;
    # each event pair costs the stream a few microseconds and delays the
    # launch after it, so the throughput pass above carries none)
    ksteps, kevery = 400, 10
    _, kstage, _ = run_plain(ea, eng, ea.RESAMPLE_SYSTEMATIC, ksteps, args.warmup, timing=1, every=kevery)
    sw_us = float(kstage[1])
    achieved = SS_BYTES_PER_LAUNCH / (sw_us * 1e-6) / 1e9
    if args.main_only:
        print(json.dumps({"ms_per_step": secs * 1e3 / args.steps, "value": value, "kernel_us": sw_us,
                          "kernel_timed_launches": ksteps // kevery, "achieved_gbs": achieved}), flush=True)
        eng.close()
        return
    # separate short pass with events around every stage (adds idle gaps;
    # breakdown only; per-stage timing runs sample+weight and sort as two launches)
    _, stage_all, _ = run_plain(ea, eng, ea.RESAMPLE_SYSTEMATIC, min(args.steps, 100), args.warmup, timing=2)
    # the same step with sample+weight and sort as two launches
    usecs, _, _ = run_plain(ea, eng, ea.RESAMPLE_SYSTEMATIC, min(args.steps, 300), args.warmup, timing=0,
                            fusion=False)
    _, ustage, _ = run_plain(ea, eng, ea.RESAMPLE_SYSTEMATIC, 100, args.warmup, timing=1, fusion=False)
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": secs * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded unicycle trajectory and range+bearing measurements)",
        "config": {"workload": "unicycle, 4096 sub-filters x 512 = 2097152 particles, 32 landmarks, "
                               "64x64 torus t=4, systematic resampling, MTGP32 seed 42",
                   "particles": N, "sub_filters": G, "particles_per_sub_filter": M, "landmarks": L,
                   "parallelism": "single GPU"},
        "settle_s": args.settle,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic_from_profiles(),
                     "traffic_source": "profiles/sample_sort_unicycle2m_traffic.json (rocprofv3 FETCH_SIZE + "
                                       "WRITE_SIZE, calibrated by a device copy)",
                     "kernel": "k_sample_sort<unicycle, SOA, m=512> (sample+weight and block sort)",
                     "kernel_us": sw_us, "kernel_timed_launches": ksteps // kevery, "bytes_per_launch": SS_BYTES_PER_LAUNCH,
                     "valu": valu_from_profiles(sw_us)},
        "stages_us": stage_dict(stage_all, fused=True),
        "two_launches": {"ms_per_step": usecs * 1e3 / min(args.steps, 300), "sample_weight_us": float(ustage[1]),
                         "note": "sample+weight and sort as separate launches (k_sample_weight, k_sort)"},
        "position_error_last_step": err,
    }
    # the same kernel where the likelihood is light: 4 landmarks at (+-10, +-10)
    # (the small unicycle workload's set) on the same 2M particles; sample+weight
    # is then HBM-bound and its HBM fraction measures the kernel's memory path
    _, stage4, _ = run_plain(ea, eng, ea.RESAMPLE_SYSTEMATIC, 100, args.warmup, timing=1, landmarks=4, box=0.0)
    _, stage4u, _ = run_plain(ea, eng, ea.RESAMPLE_SYSTEMATIC, 100, args.warmup, timing=1, landmarks=4, box=0.0,
                              fusion=False)
    a4 = SS_BYTES_PER_LAUNCH / (stage4[1] * 1e-6) / 1e9
    a4u = SW_BYTES_PER_LAUNCH / (stage4u[1] * 1e-6) / 1e9
    out["roofline"]["light_likelihood"] = {"workload": "same 2097152 particles, 4 landmarks", "kernel_us": stage4[1],
                                           "achieved": a4, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                           "frac": a4 / HBM_PEAK_GBS,
                                           "sample_weight_alone": {"kernel": "k_sample_weight<unicycle, SOA>",
                                                                   "kernel_us": float(stage4u[1]), "achieved": a4u,
                                                                   "frac": a4u / HBM_PEAK_GBS}}
    gsecs, _, _ = run_plain(ea, eng, ea.RESAMPLE_SYSTEMATIC, args.steps, args.warmup, graph=True)
    out["graph_replay"] = {"value": N * args.steps / gsecs, "ms_per_step": gsecs * 1e3 / args.steps}
    if not args.no_alias:
        asecs, astage, aerr = run_plain(ea, eng, ea.RESAMPLE_ALIAS, max(1, args.steps // 2), args.warmup, timing=2)
        out["alias"] = {"value": N * max(1, args.steps // 2) / asecs, "unit": "particles/s",
                        "ms_per_step": asecs * 1e3 / max(1, args.steps // 2),
                        "stages_us": stage_dict(astage, fused=True)}
    if not args.no_accuracy:
        out["accuracy"] = run_accuracy(ea, eng)
        out["accuracy_headline"] = run_accuracy_headline(ea, eng)
    if not args.no_dh6:
        out["dh6"] = run_dh6(ea, eng, min(args.steps, 500), args.warmup)
    if not args.no_layouts:
        out["layouts"] = run_layouts(ea, eng)
    if not args.no_batched:
        out["batched"] = run_batched(ea, eng)
    if not args.no_served:
        out["served"] = run_served(ea, eng)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
    eng.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
