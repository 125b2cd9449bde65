"""The example model randomwalk3, which the engine knows only from its line in
the model list, in the batched and served forms: every sub-filter class
(m = 64 ... 1024, fused and unfused sampling) bit-identical to the oracle,
posterior means within 1e-4 * |m64| + 1e-6 of the oracle's FP64 twin.  The
single-filter cases sit with the other models' in test_gpu_filter.py."""
import ctypes as C

import numpy as np
import pytest

import esthera_amd as ea
import oracle as O

pytestmark = pytest.mark.gpu
GiB = 1 << 30
SIZES = [(64, 5), (128, 3), (64, 3), (256, 4), (512, 2), (1024, 2), (64, 6)]


def rw3_specs(nf, steps, seed):
    """nf randomwalk3 filters of the SIZES mix; every third with the
    long-stream per-filter RNG mode."""
    specs = []
    for i in range(nf):
        m, G = SIZES[i % len(SIZES)]
        rng = ea.RNG_MTGP_PER_FILTER_LONG if i % 3 == 2 else ea.RNG_MTGP_PER_FILTER
        d = ea.filter_desc(ea.MODEL_RANDOMWALK3, m=m, G=G, t=1, topology=ea.TOPO_RING,
                           resampler=ea.RESAMPLE_SYSTEMATIC, weights=ea.WEIGHTS_LOG, rng=rng, seed=seed + i)
        p, ctrl, meas, _ = ea.randomwalk3_scenario(seed + 100 + i, steps, sigma_q=0.05 + 0.01 * i)
        specs.append((d, p, [(ctrl[k], meas[k], 0.1) for k in range(steps)]))
    return specs


def check(i, st, w, est, o):
    ost, ow = o.read(3)
    m32, m64, best = o.estimate(3)
    assert np.array_equal(st.view(np.uint32), ost.view(np.uint32)), f"filter {i} state"
    assert np.array_equal(w.view(np.uint32), ow.view(np.uint32)), f"filter {i} weights"
    assert est.best_weight == best[0] and est.best_block == int(best[1]), f"filter {i} best"
    mean = np.array(est.mean[:3])
    assert np.all(np.abs(mean - m64) <= 1e-4 * np.abs(m64) + 1e-6), f"filter {i}: {mean} vs FP64 {m64}"


def test_randomwalk3_batch_vs_oracle(engine):
    nf, steps = 10, 5
    specs = rw3_specs(nf, steps, 700)
    b = ea.Batch(engine, [s[0] for s in specs], [s[1] for s in specs])
    b.init_uniform(np.zeros((nf, 3), np.float32))
    for k in range(steps):
        ctrls = (ea.RandomWalk3Control * nf)(*[s[2][k][0] for s in specs])
        meas = (ea.RandomWalk3Measurement * nf)(*[s[2][k][1] for s in specs])
        b.step(ctrls, meas, np.full(nf, 0.1, np.float32))
    b.sync()
    for i, (d, p, inputs) in enumerate(specs):
        print(f"filter {i}: m={d.num_particles} G={d.num_blocks}", flush=True)
        o = O.Filter(d, p, C.sizeof(p))
        o.init_uniform(np.zeros(3, np.float32))
        for inp in inputs:
            o.step(*inp)
        st, w = b.particles(i)
        check(i, st, w, b.estimate(i), o)
    b.close()


@pytest.mark.parametrize("tiled", [False, True])
def test_randomwalk3_served_vs_oracle(engine, tiled):
    """Served randomwalk3 tracks against the oracle; `tiled`: every class
    fused and resampled at the other tiling (32 lanes at m = 64, two waves at
    m >= 256)."""
    steps = 4
    specs = rw3_specs(5, steps, 900)
    if tiled:
        for k in range(5):
            engine.set_batch_fusion(64 << k, ea.FUSION_ON)
            engine.set_batch_resample_lanes(64 << k, 32 if k == 0 else (128 if k >= 2 else 64))
    try:
        served_vs_oracle(engine, specs, steps)
    finally:
        for k in range(5):
            engine.set_batch_fusion(64 << k, ea.FUSION_AUTO)
            engine.set_batch_resample_lanes(64 << k, 0)


def served_vs_oracle(engine, specs, steps):
    srv = ea.Server(engine, 16, GiB)
    tids = [srv.open(d, p, np.zeros(3, np.float32))[0] for d, p, _ in specs]
    expect = {}
    for k in range(steps):
        for i, (d, p, inputs) in enumerate(specs):
            expect[srv.update(tids[i], *inputs[k])] = (i, k)
        srv.pump()
    srv.drain()
    got = {r.request_id: r for r in srv.poll(1000) if r.kind == 1}
    assert set(got) == set(expect)
    by_step = {v: rid for rid, v in expect.items()}
    for i, (d, p, inputs) in enumerate(specs):
        print(f"track {i}: m={d.num_particles} G={d.num_blocks}", flush=True)
        o = O.Filter(d, p, C.sizeof(p))
        o.init_uniform(np.zeros(3, np.float32))
        for k in range(steps):
            o.step(*inputs[k])
            r = got[by_step[(i, k)]]
            m32, m64, best = o.estimate(3)
            assert r.status == 0 and r.estimate.step == k + 1
            assert r.estimate.best_weight == best[0] and r.estimate.best_block == int(best[1])
        st, w = srv.particles(tids[i])
        check(i, st, w, got[by_step[(i, steps - 1)]].estimate, o)
    srv.close()
