"""The batched form: many filters of different sizes, landmark sets and seeds
stepped by one launch sequence per step.  Every filter's particles must be
bit-identical to the same filter stepped alone (ESTH_RNG_MTGP_PER_FILTER) and
to the oracle's run of that filter, in eager and graph-replay mode, for each
resampler and topology; posterior means within 1e-4 * |m64| + 1e-6 of the
oracle's FP64 twin."""
import ctypes as C
import os

import numpy as np
import pytest

import esthera_amd as ea
import oracle as O


def oracle_run(desc, params, inputs, init, S):
    """The oracle's run of one filter: (state, weights, m32, m64, best)."""
    o = O.Filter(desc, params, C.sizeof(params) if params is not None else 0)
    o.init_uniform(init)
    for c, m, dt in inputs:
        o.step(c, m, dt)
    st, w = o.read(S)
    m32, m64, best = o.estimate(S)
    return st, w, m32, m64, best


def check_vs_oracle(i, st, w, est, ref, S):
    ost, ow, m32, m64, best = ref
    assert np.array_equal(st.view(np.uint32), ost.view(np.uint32)), f"filter {i} state differs from the oracle"
    assert np.array_equal(w.view(np.uint32), ow.view(np.uint32)), f"filter {i} weights differ from the oracle"
    assert est.best_weight == best[0] and est.best_block == int(best[1])
    mean = np.array(est.mean[:S])
    assert np.all(np.abs(mean - m64) <= 1e-4 * np.abs(m64) + 1e-6), f"filter {i}: {mean} vs FP64 {m64}"

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def make_filters(nf, seed, sizes, topology=ea.TOPO_RING, t=1, resampler=ea.RESAMPLE_SYSTEMATIC, threshold=1.0,
                 carry=0, adaptive=0):
    rng = np.random.default_rng(seed)
    descs, params, scen = [], [], []
    for i in range(nf):
        m, G = sizes[i % len(sizes)]
        d = ea.filter_desc(ea.MODEL_UNICYCLE, m=m, G=G, t=t, topology=topology, resampler=resampler,
                           threshold=threshold, weights=ea.WEIGHTS_LOG, rng=ea.RNG_MTGP_PER_FILTER,
                           seed=int(rng.integers(1, 1 << 30)), carry=carry, adaptive=adaptive)
        p, ctrl, meas, truth = ea.unicycle_scenario(seed * 1000 + i, 8, 20.0, 8)
        descs.append(d)
        params.append(p)
        scen.append((ctrl, meas))
    return descs, params, scen


def step_inputs(scen, k):
    nf = len(scen)
    ctrls = (ea.UnicycleControl * nf)(*[s[0][k] for s in scen])
    meas = (ea.UnicycleMeasurement * nf)(*[s[1][k] for s in scen])
    return ctrls, meas, np.full(nf, 0.1, np.float32)


CASES = [
    ("ring-systematic", dict(sizes=[(64, 3), (128, 5), (256, 2), (512, 4), (1024, 1), (64, 1)])),
    ("torus-multinomial", dict(sizes=[(256, 16), (512, 4), (64, 64)], topology=ea.TOPO_TORUS, t=2,
                               resampler=ea.RESAMPLE_MULTINOMIAL)),
    ("ring-alias", dict(sizes=[(64, 4), (256, 3), (1024, 2)], resampler=ea.RESAMPLE_ALIAS)),
    ("all2all", dict(sizes=[(128, 8), (256, 5)], topology=ea.TOPO_ALL2ALL, t=2)),
    ("ring-carry-gate", dict(sizes=[(64, 3), (256, 4), (128, 2)], threshold=0.5, carry=1)),
    ("torus-alias-adaptive-carry", dict(sizes=[(256, 16), (1024, 4), (64, 64)], topology=ea.TOPO_TORUS, t=2,
                                        resampler=ea.RESAMPLE_ALIAS, threshold=0.02, adaptive=1, carry=1)),
    ("all2all-alias", dict(sizes=[(128, 8), (512, 5)], topology=ea.TOPO_ALL2ALL, t=2, resampler=ea.RESAMPLE_ALIAS)),
    # skipped sub-filters without carried weights keep their propagated state
    ("ring-gate", dict(sizes=[(128, 4), (256, 3), (64, 2), (512, 2)], threshold=0.5)),
    ("ring-alias-adaptive", dict(sizes=[(256, 4), (128, 6), (1024, 2)], resampler=ea.RESAMPLE_ALIAS, threshold=0.02,
                                 adaptive=1)),
]


@pytest.mark.parametrize("name,kw", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("graph", [False, True])
def test_batch_equals_filters_alone(name, kw, graph, engine):
    nf = 9
    descs, params, scen = make_filters(nf, 3, **kw)
    b = ea.Batch(engine, descs, params)
    b.use_graph(graph)
    b.init_uniform(np.zeros((nf, 3), np.float32))
    alone = []
    for i in range(nf):
        f = ea.Filter(engine, descs[i], params[i])
        f.init_uniform(np.zeros(3, np.float32))
        alone.append(f)
    steps = 5
    for k in range(steps):
        ctrls, meas, dts = step_inputs(scen, k)
        b.step(ctrls, meas, dts)
        for i, f in enumerate(alone):
            f.step(scen[i][0][k], scen[i][1][k], 0.1)
    for i, f in enumerate(alone):
        sb, wb = b.particles(i)
        sa, wa = f.particles()
        assert np.array_equal(sb.view(np.uint32), sa.view(np.uint32)), f"filter {i} state differs"
        assert np.array_equal(wb.view(np.uint32), wa.view(np.uint32)), f"filter {i} weights differ"
        eb, ea_ = b.estimate(i), f.estimate()
        np.testing.assert_allclose(np.array(eb.mean[:3]), np.array(ea_.mean[:3]), rtol=1e-9, atol=1e-9)
        assert eb.best_weight == ea_.best_weight and eb.best_block == ea_.best_block
        ref = oracle_run(descs[i], params[i], [(scen[i][0][k], scen[i][1][k], 0.1) for k in range(steps)],
                         np.zeros(3, np.float32), 3)
        check_vs_oracle(i, sb, wb, eb, ref, 3)
        f.close()
    t = b.step_times()
    assert len(t) == steps and np.all(t > 0)
    b.close()


def test_batch_dh6_filters_equal_alone(engine):
    """dh6 filters (12-dimensional state) in one batch, fused size classes
    (m = 128, 256) beside m = 64 and 512, with a gate that skips sub-filters."""
    sizes = [(128, 4), (256, 2), (64, 6), (512, 2)]
    nf, steps = 6, 4
    descs, params, scen, inits = [], [], [], []
    for i in range(nf):
        m, G = sizes[i % len(sizes)]
        descs.append(ea.filter_desc(ea.MODEL_DH6, m=m, G=G, t=1, topology=ea.TOPO_RING,
                                    resampler=ea.RESAMPLE_SYSTEMATIC, weights=ea.WEIGHTS_LOG, threshold=0.5,
                                    rng=ea.RNG_MTGP_PER_FILTER, seed=300 + i))
        p, meas, truth = ea.dh6_scenario(70 + i, steps)
        params.append(p)
        scen.append(meas)
        inits.append(truth[0])
    b = ea.Batch(engine, descs, params)
    b.init_uniform(np.array(inits, np.float32))
    alone = []
    for i in range(nf):
        f = ea.Filter(engine, descs[i], params[i])
        f.init_uniform(inits[i])
        alone.append(f)
    for k in range(steps):
        ctrls = (ea.DH6Control * nf)()
        meas = (ea.DH6Measurement * nf)(*[sc[k] for sc in scen])
        b.step(ctrls, meas, np.full(nf, 0.01, np.float32))
        for i, f in enumerate(alone):
            f.step(ea.DH6Control(), scen[i][k], 0.01)
    for i, f in enumerate(alone):
        sb, wb = b.particles(i)
        sa, wa = f.particles()
        assert np.array_equal(sb.view(np.uint32), sa.view(np.uint32)), f"filter {i} state differs"
        assert np.array_equal(wb.view(np.uint32), wa.view(np.uint32)), f"filter {i} weights differ"
        ref = oracle_run(descs[i], params[i], [(ea.DH6Control(), scen[i][k], 0.01) for k in range(steps)],
                         inits[i], 12)
        check_vs_oracle(i, sb, wb, b.estimate(i), ref, 12)
        f.close()
    b.close()


def test_batch_camera_filters_equal_alone(engine):
    """robotarm5_camera filters with their own calibrations in one batch."""
    sizes = [(64, 4), (256, 2), (128, 3)]
    nf, steps = 5, 4
    descs, params, scen, inits = [], [], [], []
    for i in range(nf):
        m, G = sizes[i % len(sizes)]
        descs.append(ea.filter_desc(ea.MODEL_ROBOTARM5_CAMERA, m=m, G=G, t=1, topology=ea.TOPO_RING,
                                    resampler=ea.RESAMPLE_SYSTEMATIC, weights=ea.WEIGHTS_LOG,
                                    rng=ea.RNG_MTGP_PER_FILTER, seed=100 + i))
        p = ea.camera_params_default()
        p.fc[0] *= 1.0 + 0.05 * i
        p.kc[1] *= 1.0 - 0.1 * i
        ctrl, meas, truth = ea.robotarm5_camera_scenario(50 + i, steps, 0.1, p)
        params.append(p)
        scen.append((ctrl, meas))
        inits.append(truth[0])
    b = ea.Batch(engine, descs, params)
    b.init_uniform(np.array(inits, np.float32))
    alone = []
    for i in range(nf):
        f = ea.Filter(engine, descs[i], params[i])
        f.init_uniform(inits[i])
        alone.append(f)
    for k in range(steps):
        ctrls = (ea.RobotArm5Control * nf)(*[sc[0][k] for sc in scen])
        meas = (ea.RobotArm5Measurement * nf)(*[sc[1][k] for sc in scen])
        b.step(ctrls, meas, np.full(nf, 0.1, np.float32))
        for i, f in enumerate(alone):
            f.step(scen[i][0][k], scen[i][1][k], 0.1)
    for i, f in enumerate(alone):
        sb, wb = b.particles(i)
        sa, wa = f.particles()
        assert np.array_equal(sb.view(np.uint32), sa.view(np.uint32)), f"filter {i} state differs"
        assert np.array_equal(wb.view(np.uint32), wa.view(np.uint32)), f"filter {i} weights differ"
        f.close()
    b.close()


def test_batch_info_and_memory(engine):
    descs, params, scen = make_filters(4, 5, [(256, 4), (512, 2)])
    used0, _ = engine.memory()
    b = ea.Batch(engine, descs, params)
    particles, streams, subs = b.info()
    assert particles == sum(d.num_particles * d.num_blocks for d in descs)
    assert subs == sum(d.num_blocks for d in descs)
    assert engine.memory()[0] > used0
    b.close()
    assert engine.memory()[0] == used0


def test_autotune_measures_all_sizes(engine):
    """The autotuner times the batch path per size class and sets the
    engine's fusion choice for each; a size the descriptor rejects (t = 40
    on a ring needs m > 64) costs +inf instead of failing the sweep."""
    p, ctrl, meas, truth = ea.unicycle_scenario(99, 8, 20.0, 2)
    d = ea.filter_desc(ea.MODEL_UNICYCLE, m=256, G=4, t=1, topology=ea.TOPO_RING, resampler=ea.RESAMPLE_SYSTEMATIC,
                       weights=ea.WEIGHTS_LOG, rng=ea.RNG_MTGP_PER_FILTER, seed=1)
    try:
        us = ea.autotune(engine, d, p, ctrl[0], meas[0], 0.1)
        assert us.shape == (5,) and np.all(us > 0) and np.all(np.isfinite(us))
        assert all(engine.batch_fusion(64 << k) in (ea.FUSION_ON, ea.FUSION_OFF) for k in range(5))
        assert engine.batch_resample_lanes(64) in (32, 64) and engine.batch_resample_lanes(128) == 64
        assert all(engine.batch_resample_lanes(64 << k) in (64, 128) for k in range(2, 5))
        m = ea.autotune_pick(us, 65536, 256)
        assert m in (256, 512, 1024)
        d.num_transfer = 40
        us = ea.autotune(engine, d, p, ctrl[0], meas[0], 0.1)
        assert np.isinf(us[0]) and np.all(np.isfinite(us[1:]))
    finally:
        for k in range(5):
            engine.set_batch_fusion(64 << k, ea.FUSION_AUTO)
            engine.set_batch_resample_lanes(64 << k, 0)


@pytest.mark.parametrize("model", [ea.MODEL_UNICYCLE, ea.MODEL_ROBOTARM5])
def test_batch_fused_and_unfused_classes_identical(engine, model):
    """Every size class sampled and sorted in one launch and in two, and
    resampled at 64 lanes per sub-filter and at the other tiling (32 at
    m = 64, 128 at m >= 256): the same particles, bit for bit (the autotuner
    picks any of these)."""
    sizes = [(64, 5), (128, 3), (64, 2), (256, 2), (512, 2), (1024, 1), (64, 1)]
    nf, steps = 7, 3
    descs, params, scen, inits = [], [], [], []
    meas_ra, ctrl_ra, dts_ra, _ = ea.load_robotarm_trace(os.path.join(HERE, "golden", "robotarm_input5.txt"))
    for i in range(nf):
        m, G = sizes[i]
        descs.append(ea.filter_desc(model, m=m, G=G, t=1, topology=ea.TOPO_RING, resampler=ea.RESAMPLE_SYSTEMATIC,
                                    weights=ea.WEIGHTS_LOG, rng=ea.RNG_MTGP_PER_FILTER, seed=300 + i))
        if model == ea.MODEL_UNICYCLE:
            p, c, ms, _ = ea.unicycle_scenario(40 + i, 8, 20.0, steps)
            params.append(p)
            scen.append((c, ms, [0.1] * steps))
            inits.append(np.zeros(3, np.float32))
        else:
            params.append(None)
            scen.append((ctrl_ra, meas_ra, dts_ra))
            inits.append(np.array([0.75 * np.pi] * 5 + [7.0, 0.0, 0.0, 0.0], np.float32))
    S = ea.STATE_DIM[model]
    out = {}
    try:
        for mode in (ea.FUSION_OFF, ea.FUSION_ON):
            for k in range(5):
                engine.set_batch_fusion(64 << k, mode)
                alt = 32 if k == 0 else (128 if k >= 2 else 64)
                engine.set_batch_resample_lanes(64 << k, 64 if mode == ea.FUSION_OFF else alt)
            b = ea.Batch(engine, descs, params if model == ea.MODEL_UNICYCLE else None)
            b.init_uniform(np.array(inits, np.float32))
            for k in range(steps):
                ct = ea.CONTROL_TYPE[model]
                mt = ea.MEAS_TYPE[model]
                ctrls = (ct * nf)(*[sc[0][k] for sc in scen])
                meas = (mt * nf)(*[sc[1][k] for sc in scen])
                b.step(ctrls, meas, np.array([sc[2][k] for sc in scen], np.float32))
            out[mode] = [b.particles(i) for i in range(nf)]
            b.close()
    finally:
        for k in range(5):
            engine.set_batch_fusion(64 << k, ea.FUSION_AUTO)
            engine.set_batch_resample_lanes(64 << k, 0)
    for i in range(nf):
        (sa, wa), (sb, wb) = out[ea.FUSION_OFF][i], out[ea.FUSION_ON][i]
        assert np.array_equal(sa.view(np.uint32), sb.view(np.uint32)), f"filter {i} state"
        assert np.array_equal(wa.view(np.uint32), wb.view(np.uint32)), f"filter {i} weights"
    # and the oracle agrees with both
    for i in (0, 1, 5):
        o = O.Filter(descs[i], params[i], C.sizeof(params[i]) if params[i] is not None else 0)
        o.init_uniform(inits[i])
        for k in range(steps):
            o.step(scen[i][0][k], scen[i][1][k], float(scen[i][2][k]))
        ost, ow = o.read(S)
        sa, wa = out[ea.FUSION_ON][i]
        assert np.array_equal(sa.view(np.uint32), ost.view(np.uint32)), f"filter {i} vs oracle"
        assert np.array_equal(wa.view(np.uint32), ow.view(np.uint32)), f"filter {i} weights vs oracle"


def test_many_filters_parallel_frames(engine):
    """Enough filters that the step's frames are built by several host threads."""
    nf = 700
    rng = np.random.default_rng(8)
    scen = [ea.unicycle_scenario(500 + j, 6, 15.0, 4) for j in range(7)]
    descs, params = [], []
    for i in range(nf):
        G = int(rng.integers(2, 5))
        descs.append(ea.filter_desc(ea.MODEL_UNICYCLE, m=64, G=G, t=1, topology=ea.TOPO_RING,
                                    resampler=ea.RESAMPLE_SYSTEMATIC, weights=ea.WEIGHTS_LOG,
                                    rng=ea.RNG_MTGP_PER_FILTER, seed=7000 + i))
        params.append(scen[i % 7][0])
    b = ea.Batch(engine, descs, params)
    b.init_uniform(np.zeros((nf, 3), np.float32))
    for k in range(3):
        ctrls = (ea.UnicycleControl * nf)(*[scen[i % 7][1][k] for i in range(nf)])
        meas = (ea.UnicycleMeasurement * nf)(*[scen[i % 7][2][k] for i in range(nf)])
        b.step(ctrls, meas, np.full(nf, 0.1, np.float32))
    for i in (0, 255, 256, 511, 699):
        f = ea.Filter(engine, descs[i], params[i])
        f.init_uniform(np.zeros(3, np.float32))
        for k in range(3):
            f.step(scen[i % 7][1][k], scen[i % 7][2][k], 0.1)
        sa, wa = f.particles()
        sb, wb = b.particles(i)
        assert np.array_equal(sa.view(np.uint32), sb.view(np.uint32)), f"filter {i}"
        assert np.array_equal(wa.view(np.uint32), wb.view(np.uint32)), f"filter {i}"
        f.close()
    b.close()
