"""The benchmark configurations at their own sizes against the oracle:

* the 6-DOF DH arm, 2048 sub-filters x 512 on the 32 x 64 torus, t = 1
  (seed 7), two whole steps bit for bit;
* the 2M-particle unicycle with the Vose alias resampler (4096 x 512, 64 x 64
  torus, t = 4, 32 landmarks, seed 42), two whole steps bit for bit;
* the 10,000-filter batch (particle counts from seed 99): every filter's
  estimate checked by properties, a handful of filters against the oracle;
* served tracks of up to 262,144 particles against the oracle.

The posterior mean is also compared with the oracle's FP64 twin (the same
prior and normals propagated in double): per component within
1e-4 * |m64| + 1e-6, i.e. relative wherever the component is not tiny."""
import ctypes
import os

import numpy as np
import pytest

import esthera_amd as ea
import oracle as O
from esthera_amd import workloads

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
GiB = 1 << 30


def posterior_close(mean, m64, rtol=1e-4, atol=1e-6):
    mean, m64 = np.asarray(mean, np.float64), np.asarray(m64, np.float64)
    return bool(np.all(np.abs(mean - m64) <= rtol * np.abs(m64) + atol)), np.abs(mean - m64)


def same_bits(a, b):
    return np.array_equal(np.asarray(a, np.float32).view(np.uint32), np.asarray(b, np.float32).view(np.uint32))


def run_vs_oracle(engine, desc, params, init, inputs, S, fused=None):
    f = ea.Filter(engine, desc, params)
    if fused is not None:
        f.set_fusion(fused)
    o = O.Filter(desc, params, ctypes.sizeof(params) if params is not None else 0)
    f.init_uniform(init)
    o.init_uniform(init)
    for k, (c, m, dt) in enumerate(inputs):
        print(f"step {k}", flush=True)
        f.step(c, m, dt)
        o.step(c, m, dt)
        gs, gw = f.particles()
        os_, ow = o.read(S)
        assert same_bits(gs, os_), f"state differs at step {k}"
        assert same_bits(gw, ow), f"weights differ at step {k}"
        est = f.estimate()
        m32, m64, best = o.estimate(S)
        assert est.best_weight == best[0] and est.best_block == int(best[1])
        assert same_bits(np.array(est.best[:S]), o.best_particle(S))
        np.testing.assert_allclose(np.array(est.mean[:S]), m32, rtol=1e-6, atol=1e-6)
        ok, err = posterior_close(est.mean[:S], m64)
        assert ok, f"step {k}: posterior mean vs FP64 twin {err}"
    f.close()


def test_dh6_2048x512_torus_bit_exact_vs_oracle(engine):
    """The robot-arm configuration at its size: 6-joint DH arm (seed 7), 12-dim
    state, 2048 sub-filters x 512 = 1,048,576 particles, 32 x 64 torus
    exchanging one particle per neighbour, systematic resampling."""
    desc = ea.filter_desc(ea.MODEL_DH6, m=512, G=2048, t=1, topology=ea.TOPO_TORUS,
                          resampler=ea.RESAMPLE_SYSTEMATIC, weights=ea.WEIGHTS_LOG, seed=7)
    p, meas, truth = ea.dh6_scenario(7, 2)
    inputs = [(ea.DH6Control(), meas[k], 0.01) for k in range(2)]
    run_vs_oracle(engine, desc, p, truth[0], inputs, 12)


def test_unicycle_alias_4096x512_bit_exact_vs_oracle(engine):
    """The 2M-particle unicycle with the Vose alias resampler: 64 x 64 torus,
    t = 4, 32 landmarks, MTGP32 seed 42 (the reference's 96-stream buffer)."""
    desc = ea.filter_desc(ea.MODEL_UNICYCLE, m=512, G=4096, t=4, topology=ea.TOPO_TORUS,
                          resampler=ea.RESAMPLE_ALIAS, weights=ea.WEIGHTS_LOG, seed=42)
    p, ctrl, meas, truth = ea.unicycle_scenario(42, 32, 50.0, 2)
    inputs = [(ctrl[k], meas[k], 0.1) for k in range(2)]
    run_vs_oracle(engine, desc, p, np.zeros(3, np.float32), inputs, 3, fused=True)


def test_batched_10k_filters_vs_oracle(engine):
    """The 10,000-filter batch (118.8 M particles) for two steps: every
    filter's estimate is finite, counts two steps and sits near its own
    scenario's truth; the smallest and largest filters and three others are
    bit-identical to the oracle, posterior means within the FP64 bar."""
    d0 = ea.filter_desc(ea.MODEL_UNICYCLE, m=256, G=4, t=1, topology=ea.TOPO_RING,
                        resampler=ea.RESAMPLE_SYSTEMATIC, weights=ea.WEIGHTS_LOG, rng=ea.RNG_MTGP_PER_FILTER_LONG, seed=1)
    scen = workloads.batched_scenarios()
    p0, c0, m0, _ = scen[0]
    us = ea.autotune(engine, d0, p0, c0[0], m0[0], 0.1)
    counts, descs, params, scen = workloads.batched_filters(us, scen=scen)
    nf = len(descs)
    assert nf == 10000 and int(counts.sum()) == 118794240
    b = ea.Batch(engine, descs, params)
    b.init_uniform(np.zeros((nf, 3), np.float32))
    steps = 2
    for k in range(steps):
        ctrls, meas = workloads.batched_inputs(scen, nf, k)
        b.step(ctrls, meas, np.full(nf, 0.1, np.float32))
    b.sync()
    print("batch stepped", flush=True)
    far = 0
    for i in range(nf):
        e = b.estimate(i)
        mean = np.array(e.mean[:3])
        assert e.step == steps and np.all(np.isfinite(mean)) and np.isfinite(e.best_weight)
        tr = scen[i % len(scen)][3][steps - 1]
        far += int(np.hypot(mean[0] - tr[0], mean[1] - tr[1]) > 0.5)
    assert far < nf // 100, far
    rng = np.random.default_rng(99)
    pick = sorted({int(np.argmin(counts)), int(np.argmax(counts)), *rng.integers(0, nf, 3).tolist()})
    for i in pick:
        print(f"filter {i}: {counts[i]} particles, m={descs[i].num_particles}", flush=True)
        o = O.Filter(descs[i], params[i], ctypes.sizeof(params[i]))
        o.init_uniform(np.zeros(3, np.float32))
        for k in range(steps):
            sc = scen[i % len(scen)]
            o.step(sc[1][k], sc[2][k], 0.1)
        gs, gw = b.particles(i)
        os_, ow = o.read(3)
        assert same_bits(gs, os_), f"filter {i} state"
        assert same_bits(gw, ow), f"filter {i} weights"
        e = b.estimate(i)
        m32, m64, best = o.estimate(3)
        assert e.best_weight == best[0] and e.best_block == int(best[1])
        ok, err = posterior_close(e.mean[:3], m64)
        assert ok, f"filter {i}: {err}"
    b.close()


def test_served_tracks_up_to_262144_vs_oracle(engine):
    """Served tracks at the top of the served size range (262,144 particles,
    unicycle and robot arm) beside small ones, opened and updated while earlier
    waves run: every update's best particle and the final particles equal the
    oracle's, posterior means within the FP64 bar."""
    meas_ra, ctrl_ra, dts_ra, _ = ea.load_robotarm_trace(os.path.join(HERE, "golden", "robotarm_input5.txt"))
    ra_init = np.array([0.75 * np.pi] * 5 + [7.0, 0.0, 0.0, 0.0], np.float32)
    specs = []
    for seed, model, m, G in [(41, ea.MODEL_UNICYCLE, 256, 1024), (42, ea.MODEL_ROBOTARM5, 512, 512),
                              (43, ea.MODEL_UNICYCLE, 64, 16), (44, ea.MODEL_UNICYCLE, 512, 512),
                              (45, ea.MODEL_ROBOTARM5, 128, 8)]:
        d = ea.filter_desc(model, m=m, G=G, t=1, topology=ea.TOPO_TORUS, resampler=ea.RESAMPLE_SYSTEMATIC,
                           weights=ea.WEIGHTS_LOG, rng=ea.RNG_MTGP_PER_FILTER, seed=seed)
        if model == ea.MODEL_UNICYCLE:
            p, c, ms, _ = ea.unicycle_scenario(seed, 8, 20.0, 4)
            specs.append((d, p, [(c[k], ms[k], 0.1) for k in range(4)], np.zeros(3, np.float32), 3))
        else:
            specs.append((d, None, [(ctrl_ra[k], meas_ra[k], float(dts_ra[k])) for k in range(4)], ra_init, 9))
    srv = ea.Server(engine, 16, 4 * GiB)
    tids, expect = [], {}
    for i, (d, p, inputs, init, S) in enumerate(specs):
        tids.append(srv.open(d, p, init)[0])
        expect[srv.update(tids[i], *inputs[0])] = (i, 0)
        srv.pump()
    for k in range(1, 4):
        for i, (d, p, inputs, init, S) in enumerate(specs):
            expect[srv.update(tids[i], *inputs[k])] = (i, k)
        srv.pump()
    srv.drain()
    got = {r.request_id: r for r in srv.poll(1000) if r.kind == 1}
    assert set(got) == set(expect)
    for i, (d, p, inputs, init, S) in enumerate(specs):
        print(f"track {i}: {d.num_particles * d.num_blocks} particles", flush=True)
        o = O.Filter(d, p, ctypes.sizeof(p) if p is not None else 0)
        o.init_uniform(init)
        for k in range(4):
            o.step(*inputs[k])
            r = next(r for rid, r in got.items() if expect[rid] == (i, k))
            m32, m64, best = o.estimate(S)
            assert r.status == 0 and r.estimate.step == k + 1
            assert r.estimate.best_weight == best[0] and r.estimate.best_block == int(best[1])
            assert same_bits(np.array(r.estimate.best[:S]), o.best_particle(S)), f"track {i} step {k}"
            ok, err = posterior_close(r.estimate.mean[:S], m64)
            assert ok, f"track {i} step {k}: {err}"
        st, w = srv.particles(tids[i])
        os_, ow = o.read(S)
        assert same_bits(st, os_) and same_bits(w, ow), f"track {i} particles"
    srv.close()
