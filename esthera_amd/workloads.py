"""Seeded workloads shared by bench.py and the GPU tests.

batched_10k: 10,000 independent unicycle filters, each with its own set of 8
landmarks (64 seeded sets), particle counts log-uniform in [256, 65536]
rounded to multiples of 64 (seed 99), ring exchange t = 1, systematic
resampling, log weights, per-filter MTGP32 streams of about 16 rounds each
(ESTH_RNG_MTGP_PER_FILTER_LONG: the throughput form); the sub-filter size of
each filter comes from the autotuner's per-size costs.
"""
import numpy as np

from . import (MODEL_UNICYCLE, RESAMPLE_SYSTEMATIC, RNG_MTGP_PER_FILTER_LONG, TOPO_RING, WEIGHTS_LOG, UnicycleControl,
               UnicycleMeasurement, autotune_pick, filter_desc, unicycle_scenario)

BATCH_SCENARIOS = 64
BATCH_SCENARIO_STEPS = 8


def batched_counts(nf=10000, seed=99):
    """Particle counts of the batched workload: log-uniform in [256, 65536],
    rounded to multiples of 64."""
    rng = np.random.default_rng(seed)
    c = np.round(np.exp(rng.uniform(np.log(256), np.log(65536), nf)) / 64) * 64
    return np.maximum(c.astype(np.int64), 256)


def batched_scenarios():
    return [unicycle_scenario(1000 + j, 8, 20.0, BATCH_SCENARIO_STEPS) for j in range(BATCH_SCENARIOS)]


def batched_filters(us_per_mparticle, nf=10000, seed=99, scen=None, sub_filter_size=None):
    """Descriptors and model parameters of the batched workload.  Each
    filter's sub-filter size is autotune_pick over `us_per_mparticle`
    (or sub_filter_size(count) when given), halved while the filter would
    otherwise have a single sub-filter (so each filter exchanges between at
    least two)."""
    counts = batched_counts(nf, seed)
    scen = scen if scen is not None else batched_scenarios()
    descs, params = [], []
    for i, n in enumerate(counts):
        m = sub_filter_size(int(n)) if sub_filter_size else autotune_pick(us_per_mparticle, int(n), 64)
        if int(n) // m < 2:
            m //= 2
        descs.append(filter_desc(MODEL_UNICYCLE, m=m, G=int(n) // m, t=1, topology=TOPO_RING,
                                 resampler=RESAMPLE_SYSTEMATIC, weights=WEIGHTS_LOG, rng=RNG_MTGP_PER_FILTER_LONG,
                                 seed=seed * 1000003 + i))
        params.append(scen[i % len(scen)][0])
    return counts, descs, params, scen


def batched_inputs(scen, nf, k):
    """Controls and measurements of step k for all filters (filter i replays
    scenario i mod 64, step k mod 8)."""
    ns, ks = len(scen), k % BATCH_SCENARIO_STEPS
    ctrls = (UnicycleControl * nf)(*[scen[i % ns][1][ks] for i in range(nf)])
    meas = (UnicycleMeasurement * nf)(*[scen[i % ns][2][ks] for i in range(nf)])
    return ctrls, meas
