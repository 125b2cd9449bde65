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
# sample+weight algorithmic bytes per particle: state read 12 + uniforms read 12
# + state write 12 + log-weight write 4
SW_BYTES_PER_PARTICLE = 40
M, G, T, L = 512, 4096, 4, 32
STAGES = ["rng_wait", "sample_weight", "sort", "argmax", "exchange_or_topt", "resample"]


def run_plain(ea, eng, resampler, steps, warmup, graph=False, timing=1):
    desc = ea.filter_desc(ea.MODEL_UNICYCLE, m=M, G=G, t=T, topology=ea.TOPO_TORUS, resampler=resampler,
                          weights=ea.WEIGHTS_LOG, rng=ea.RNG_MTGP_COMPAT, seed=42)
    params, ctrl, meas, truth = ea.unicycle_scenario(42, L, 50.0, warmup + steps)
    f = ea.Filter(eng, desc, params)
    import numpy as np
    f.init_uniform(np.zeros(3, np.float32))
    if graph:
        f.use_graph(True)
    for k in range(warmup):
        f.step(ctrl[k], meas[k], 0.1)
    ea.device_sync()
    if not graph:
        f.set_timing(timing)
    t0 = time.perf_counter()
    for k in range(warmup, warmup + steps):
        f.step(ctrl[k], meas[k], 0.1)
    ea.device_sync()
    t1 = time.perf_counter()
    stage = [float(x) for x in f.stage_times()] if not graph else [float("nan")] * 6
    est = f.estimate()
    k = warmup + steps - 1
    err = float(np.hypot(est.mean[0] - truth[k, 0], est.mean[1] - truth[k, 1]))
    f.close()
    return t1 - t0, stage, err


def cpu_baseline(budget_s=15.0, max_steps=200):
    """The oracle's FP32 restatement of the reference step (the same MTGP
    stream, bitonic network, scan and resampler), OpenMP over sub-filters, on
    a bounded sample of the workload: the full 4096 x 512 filter for as many
    steps as fit in about `budget_s` seconds."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import numpy as np

    import esthera_amd as ea
    import oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    oracle.lib().or_set_threads(threads)
    Gs = G
    desc = ea.filter_desc(ea.MODEL_UNICYCLE, m=M, G=Gs, t=T, topology=ea.TOPO_TORUS,
                          resampler=ea.RESAMPLE_SYSTEMATIC, weights=ea.WEIGHTS_LOG, rng=ea.RNG_MTGP_COMPAT, seed=42)
    params, ctrl, meas, truth = ea.unicycle_scenario(42, L, 50.0, max_steps + 1)
    o = oracle.Filter(desc, params, ctypes.sizeof(params))
    o.set_f64(False)
    o.init_uniform(np.zeros(3, np.float32))
    o.step(ctrl[0], meas[0], 0.1)
    n = 0
    t0 = time.perf_counter()
    while n < max_steps and (time.perf_counter() - t0) < budget_s:
        o.step(ctrl[n + 1], meas[n + 1], 0.1)
        n += 1
    dt = time.perf_counter() - t0
    return {"value": M * Gs * n / dt, "unit": "particles/s", "cores": threads, "kind": "port",
            "sample": f"{n} steps (after 1 untimed) of the full 4096 x 512 workload, FP32 oracle "
                      f"restatement, OpenMP {threads} threads, {dt:.1f} s"}


def traffic_from_profiles():
    p = os.path.join(REPO, "profiles", "sample_weight_c1_traffic.json")
    if not os.path.exists(p):
        return None
    with open(p) as fh:
        return json.load(fh).get("bytes_per_launch")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-alias", action="store_true", help="skip the Vose alias variant")
    ap.add_argument("--graph-only", action="store_true", help="only the HIP-graph replay run (profiling)")
    ap.add_argument("--copy-calib", type=int, default=0,
                    help="also run one device copy of this many MiB (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              f"(the multi-GPU sharding is described in DESIGN.md)", file=sys.stderr)
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
    # timed region: eager steps, hipEvents bracketing only the sample+weight kernel
    secs, stage, err = run_plain(ea, eng, ea.RESAMPLE_SYSTEMATIC, args.steps, args.warmup)
    value = N * args.steps / secs
    # separate short pass with events around every stage (adds idle gaps; breakdown only)
    _, stage_all, _ = run_plain(ea, eng, ea.RESAMPLE_SYSTEMATIC, min(args.steps, 100), args.warmup, timing=2)
    sw_us = stage[1]
    achieved = SW_BYTES_PER_PARTICLE * N / (sw_us * 1e-6) / 1e9
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
        "config": {"workload": "unicycle c1: 4096 sub-filters x 512 = 2097152 particles, 32 landmarks, "
                               "64x64 torus t=4, systematic resampling, MTGP32 seed 42",
                   "particles": N, "sub_filters": G, "particles_per_sub_filter": M, "landmarks": L,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic_from_profiles(),
                     "kernel": "k_sample_weight<unicycle>", "kernel_us": sw_us,
                     "bytes_per_launch": SW_BYTES_PER_PARTICLE * N},
        "stages_us": {k: v for k, v in zip(STAGES, stage_all) if v == v},
        "position_error_last_step": err,
    }
    gsecs, _, _ = run_plain(ea, eng, ea.RESAMPLE_SYSTEMATIC, args.steps, args.warmup, graph=True)
    out["graph_replay"] = {"value": N * args.steps / gsecs, "ms_per_step": gsecs * 1e3 / args.steps}
    if not args.no_alias:
        asecs, astage, aerr = run_plain(ea, eng, ea.RESAMPLE_ALIAS, max(1, args.steps // 2), args.warmup, timing=2)
        out["alias"] = {"value": N * max(1, args.steps // 2) / asecs, "unit": "particles/s",
                        "ms_per_step": asecs * 1e3 / max(1, args.steps // 2),
                        "stages_us": {k: v for k, v in zip(STAGES, astage) if v == v}}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
    eng.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
