"""Whole-step parity: the engine's filter step (all kernels, fused Box-Muller,
device-resident state) against the oracle's restatement of the reference's
step loop (pf_generic.c:438-808), on the same seed and inputs.  Particle
states after every step are compared bit for bit; posterior means against the
FP64 twin per component within 1e-4 * |m64| + 1e-6 (relative wherever the
component is not tiny).  The camera model's pixel likelihood (1 px noise at a
focal length of ~400 px) turns FP32 rounding in the projection into log-weight
differences of ~1e-3, so its means are held to 1e-4 * |m64| + 1e-4; the
oracle's own FP32 path shows the same gap to its FP64 twin."""
import os

import numpy as np
import pytest

import esthera_amd as ea
import oracle as O

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def make_case(model, steps):
    if model == ea.MODEL_UNICYCLE:
        p, ctrl, meas, truth = ea.unicycle_scenario(1234, 4, 0.0, steps)
        return p, ctrl, meas, np.zeros(3, np.float32), truth, 0.1
    if model == ea.MODEL_DH6:
        p, meas, truth = ea.dh6_scenario(7, steps)
        ctrl = [ea.DH6Control() for _ in range(steps)]
        return p, ctrl, meas, truth[0].copy(), truth[1:], 0.01
    if model == ea.MODEL_UNICYCLE10:
        meas, ctrl, dts, truth = ea.load_unicycle10_trace(os.path.join(HERE, "golden", "unicycle_input.txt"))
        init = np.array([20.0, 40.0, -np.pi / 2], np.float32)
        return ea.unicycle10_params_default(), ctrl, meas, init, truth, dts
    if model == ea.MODEL_RANDOMWALK3:
        p, ctrl, meas, truth = ea.randomwalk3_scenario(31, steps)
        return p, ctrl, meas, np.zeros(3, np.float32), truth, 0.1
    if model == ea.MODEL_ROBOTARM5_CAMERA:
        ctrl, meas, truth = ea.robotarm5_camera_scenario(21, steps, 0.1)
        return ea.camera_params_default(), ctrl, meas, truth[0].copy(), truth[1:], 0.1
    meas, ctrl, dts, truth = ea.load_robotarm_trace(os.path.join(HERE, "golden", "robotarm_input5.txt"))
    a0 = np.float32(0.75 * np.pi)
    init = np.array([a0] * 5 + [7.0, 0.0, 0.0, 0.0], np.float32)
    return None, ctrl, meas, init, truth, dts


def run_pair(engine, desc, steps, check_every=1):
    model = desc.model
    params, ctrl, meas, init, truth, dt = make_case(model, steps)
    f = ea.Filter(engine, desc, params)
    f.init_uniform(init)
    o = O.Filter(desc, params, 0 if params is None else __import__("ctypes").sizeof(params))
    o.init_uniform(init)
    S = ea.STATE_DIM[model]
    worst = 0.0
    for k in range(steps):
        d = float(dt[k]) if np.ndim(dt) else dt
        f.step(ctrl[k], meas[k], d)
        o.step(ctrl[k], meas[k], d)
        if k % check_every == 0 or k == steps - 1:
            gst, gw = f.particles()
            ost, ow = o.read(S)
            assert np.array_equal(gst.view(np.uint32), ost.view(np.uint32)), f"state differs at step {k}"
            assert np.array_equal(gw.view(np.uint32), ow.view(np.uint32)), f"weights differ at step {k}"
            est = f.estimate()
            m32, m64, best = o.estimate(S)
            mean = np.array(est.mean[:S])
            atol = 1e-4 if model == ea.MODEL_ROBOTARM5_CAMERA else 1e-6
            worst = max(worst, float(np.max(np.abs(mean - m64) / (1e-4 * np.abs(m64) + atol))))
            np.testing.assert_allclose(mean, m32, rtol=1e-6, atol=1e-6)
            assert est.best_weight == best[0] and est.best_block == int(best[1])
    return worst


CASES = [
    ("unicycle-multinomial-ring-log", dict(model=ea.MODEL_UNICYCLE, m=256, G=16, t=1, weights=ea.WEIGHTS_LOG)),
    ("unicycle-systematic-torus-log", dict(model=ea.MODEL_UNICYCLE, m=256, G=16, t=1, topology=ea.TOPO_TORUS,
                                           resampler=ea.RESAMPLE_SYSTEMATIC, weights=ea.WEIGHTS_LOG)),
    ("unicycle-alias-torus-log", dict(model=ea.MODEL_UNICYCLE, m=512, G=16, t=4, topology=ea.TOPO_TORUS,
                                      resampler=ea.RESAMPLE_ALIAS, weights=ea.WEIGHTS_LOG)),
    ("unicycle-all2all-gate", dict(model=ea.MODEL_UNICYCLE, m=128, G=32, t=2, topology=ea.TOPO_ALL2ALL,
                                   threshold=0.5, weights=ea.WEIGHTS_LOG)),
    ("robotarm-reference-linear", dict(model=ea.MODEL_ROBOTARM5, m=64, G=16, t=1)),
    ("robotarm-alias-adaptive", dict(model=ea.MODEL_ROBOTARM5, m=256, G=8, t=1, resampler=ea.RESAMPLE_ALIAS,
                                     adaptive=1, threshold=0.01)),
    ("dh6-torus-log", dict(model=ea.MODEL_DH6, m=512, G=8, t=1, topology=ea.TOPO_TORUS, weights=ea.WEIGHTS_LOG)),
    # scan resampling with two waves per sub-filter (m >= 512)
    ("unicycle-m512-systematic-torus-t4", dict(model=ea.MODEL_UNICYCLE, m=512, G=16, t=4, topology=ea.TOPO_TORUS,
                                                resampler=ea.RESAMPLE_SYSTEMATIC, weights=ea.WEIGHTS_LOG)),
    ("unicycle-m1024-multinomial-gate", dict(model=ea.MODEL_UNICYCLE, m=1024, G=8, t=2, threshold=0.5,
                                             weights=ea.WEIGHTS_LOG)),
    ("robotarm-m512-linear-all2all", dict(model=ea.MODEL_ROBOTARM5, m=512, G=8, t=2, topology=ea.TOPO_ALL2ALL)),
    ("unicycle10-reference-linear", dict(model=ea.MODEL_UNICYCLE10, m=128, G=16, t=1)),
    ("unicycle-carry-gate-log", dict(model=ea.MODEL_UNICYCLE, m=256, G=16, t=1, topology=ea.TOPO_TORUS,
                                     resampler=ea.RESAMPLE_SYSTEMATIC, threshold=0.5, weights=ea.WEIGHTS_LOG,
                                     carry=1)),
    ("robotarm-carry-gate-linear", dict(model=ea.MODEL_ROBOTARM5, m=128, G=8, t=1, threshold=0.4, carry=1)),
    ("unicycle-alias-adaptive-carry", dict(model=ea.MODEL_UNICYCLE, m=256, G=8, t=2, resampler=ea.RESAMPLE_ALIAS,
                                           adaptive=1, threshold=0.02, weights=ea.WEIGHTS_LOG, carry=1)),
    ("unicycle10-alias-log", dict(model=ea.MODEL_UNICYCLE10, m=256, G=8, t=2, resampler=ea.RESAMPLE_ALIAS,
                                  weights=ea.WEIGHTS_LOG)),
    ("camera-ring-log", dict(model=ea.MODEL_ROBOTARM5_CAMERA, m=256, G=16, t=1, weights=ea.WEIGHTS_LOG)),
    ("camera-reference-linear", dict(model=ea.MODEL_ROBOTARM5_CAMERA, m=64, G=16, t=1)),
    ("camera-alias-log", dict(model=ea.MODEL_ROBOTARM5_CAMERA, m=256, G=8, t=2, resampler=ea.RESAMPLE_ALIAS,
                              weights=ea.WEIGHTS_LOG)),
    ("camera-torus-carry-gate", dict(model=ea.MODEL_ROBOTARM5_CAMERA, m=512, G=16, t=4, topology=ea.TOPO_TORUS,
                                     resampler=ea.RESAMPLE_SYSTEMATIC, weights=ea.WEIGHTS_LOG, carry=1,
                                     threshold=0.5)),
    ("unicycle-alias-all2all", dict(model=ea.MODEL_UNICYCLE, m=128, G=32, t=2, topology=ea.TOPO_ALL2ALL,
                                    resampler=ea.RESAMPLE_ALIAS, weights=ea.WEIGHTS_LOG)),
    ("dh6-alias-torus-m1024", dict(model=ea.MODEL_DH6, m=1024, G=8, t=4, topology=ea.TOPO_TORUS,
                                   resampler=ea.RESAMPLE_ALIAS, weights=ea.WEIGHTS_LOG, threshold=0.7, carry=1)),
    ("robotarm-alias-m32", dict(model=ea.MODEL_ROBOTARM5, m=32, G=16, t=2, resampler=ea.RESAMPLE_ALIAS)),
    # the example model, picked up from the model list alone
    ("randomwalk3-systematic-torus-log", dict(model=ea.MODEL_RANDOMWALK3, m=512, G=16, t=4, topology=ea.TOPO_TORUS,
                                              resampler=ea.RESAMPLE_SYSTEMATIC, weights=ea.WEIGHTS_LOG)),
    ("randomwalk3-reference-linear", dict(model=ea.MODEL_RANDOMWALK3, m=64, G=16, t=1)),
    ("randomwalk3-alias-carry-gate", dict(model=ea.MODEL_RANDOMWALK3, m=256, G=8, t=2, resampler=ea.RESAMPLE_ALIAS,
                                          weights=ea.WEIGHTS_LOG, carry=1, threshold=0.5)),
    ("unicycle-per-filter-long-rng", dict(model=ea.MODEL_UNICYCLE, m=512, G=24, t=2, topology=ea.TOPO_TORUS,
                                          resampler=ea.RESAMPLE_SYSTEMATIC, weights=ea.WEIGHTS_LOG,
                                          rng=ea.RNG_MTGP_PER_FILTER_LONG)),
    ("unicycle-small-m", dict(model=ea.MODEL_UNICYCLE, m=16, G=12, t=2, weights=ea.WEIGHTS_LOG,
                              rng=ea.RNG_MTGP_PER_FILTER)),
]


@pytest.mark.parametrize("name,kw", CASES, ids=[c[0] for c in CASES])
def test_filter_steps_bit_exact(name, kw, engine):
    desc = ea.filter_desc(seed=1234, **kw)
    worst = run_pair(engine, desc, 20)
    assert worst < 1.0, f"posterior mean vs FP64 oracle: {worst:.3f} of the bar"


def test_graph_replay_equals_eager(engine):
    desc = ea.filter_desc(ea.MODEL_UNICYCLE, m=512, G=64, t=4, topology=ea.TOPO_TORUS, weights=ea.WEIGHTS_LOG,
                          seed=42)
    params, ctrl, meas, init, truth, dt = make_case(ea.MODEL_UNICYCLE, 12)
    a, b = ea.Filter(engine, desc, params), ea.Filter(engine, desc, params)
    a.init_uniform(init)
    b.init_uniform(init)
    b.use_graph(True)
    for k in range(12):
        a.step(ctrl[k], meas[k], dt)
        b.step(ctrl[k], meas[k], dt)
        if k % 3 == 2:
            sa, wa = a.particles()
            sb, wb = b.particles()
            assert np.array_equal(sa.view(np.uint32), sb.view(np.uint32))
            assert np.array_equal(wa.view(np.uint32), wb.view(np.uint32))


def test_snapshot_restore_resumes_identically(engine):
    desc = ea.filter_desc(ea.MODEL_UNICYCLE, m=256, G=16, t=1, weights=ea.WEIGHTS_LOG, seed=9)
    params, ctrl, meas, init, truth, dt = make_case(ea.MODEL_UNICYCLE, 10)
    a = ea.Filter(engine, desc, params)
    a.init_uniform(init)
    for k in range(4):
        a.step(ctrl[k], meas[k], dt)
    snap = a.snapshot()
    for k in range(4, 10):
        a.step(ctrl[k], meas[k], dt)
    ref, _ = a.particles()
    b = ea.Filter(engine, desc, params)
    b.restore(snap)
    for k in range(4, 10):
        b.step(ctrl[k], meas[k], dt)
    got, _ = b.particles()
    assert np.array_equal(ref.view(np.uint32), got.view(np.uint32))


def test_reference_driver_robotarm_input5(engine):
    """esth_particle_filter(64, 16, 1, 1.0, input5), the drop-in for
    particle_filter(), against the oracle's run of the same loop
    (pf_generic.c:634-696,814): the reference's defaults (ring, multinomial,
    linear weights, seed 1234), particles starting at pf_generic.c:346-361,
    estimate_error of the best particle each step (pf_model_host.h:56-69),
    accumulated in double and averaged over the 100 rows."""
    import ctypes
    path = os.path.join(HERE, "golden", "robotarm_input5.txt")
    err, steps = ea.particle_filter(64, 16, 1, 1.0, path)
    assert steps == 100
    meas, ctrl, dts, truth = ea.load_robotarm_trace(path)
    desc = ea.filter_desc(ea.MODEL_ROBOTARM5, m=64, G=16, t=1, topology=ea.TOPO_RING,
                          resampler=ea.RESAMPLE_MULTINOMIAL, threshold=1.0, seed=1234)
    o = O.Filter(desc, None, 0)
    a0 = np.float32(0.75 * np.pi)
    o.init_uniform(np.array([a0] * 5 + [7.0, 0.0, 0.0, 0.0], np.float32))
    total = 0.0
    for k in range(steps):
        o.step(ctrl[k], meas[k], float(dts[k]))
        best = o.best_particle(9)
        best[7:] = 0.0
        total += ea.lib().esth_robotarm5_estimate_error(ea.ptr(best), ea.ptr(np.ascontiguousarray(truth[k])))
    assert err == total / steps, (err, total / steps)


def test_reference_unicycle_driver(engine):
    """The reference unicycle program's loop over its own input file: the best
    particle follows the recorded truth."""
    err, steps = ea.unicycle10_filter(256, 16, 1, 1.0, os.path.join(HERE, "golden", "unicycle_input.txt"))
    assert steps == 131
    assert np.isfinite(err) and err < 1.5, err


def test_unicycle_tracks_truth(engine):
    desc = ea.filter_desc(ea.MODEL_UNICYCLE, m=256, G=16, t=1, topology=ea.TOPO_TORUS, weights=ea.WEIGHTS_LOG,
                          seed=1234)
    params, ctrl, meas, init, truth, dt = make_case(ea.MODEL_UNICYCLE, 100)
    f = ea.Filter(engine, desc, params)
    f.init_uniform(init)
    err = []
    for k in range(100):
        f.step(ctrl[k], meas[k], dt)
        est = f.estimate()
        err.append(np.hypot(est.mean[0] - truth[k, 0], est.mean[1] - truth[k, 1]))
    rmse = float(np.sqrt(np.mean(np.square(err))))
    assert rmse < 0.2, rmse


def test_two_million_particle_properties(engine):
    """The 2M-particle configuration (64x64 torus of 512): one step, checked
    through properties that do not need the oracle at this size."""
    desc = ea.filter_desc(ea.MODEL_UNICYCLE, m=512, G=4096, t=4, topology=ea.TOPO_TORUS, weights=ea.WEIGHTS_LOG,
                          seed=42)
    params, ctrl, meas, truth = ea.unicycle_scenario(42, 32, 50.0, 3)
    f = ea.Filter(engine, desc, params)
    f.init_uniform(np.zeros(3, np.float32))
    for k in range(3):
        f.step(ctrl[k], meas[k], 0.1)
    st, w = f.particles()
    assert np.all(np.isfinite(st))
    w = w.reshape(4096, 512)
    head = w[:, : 512 - 16]
    assert np.all(head[:, :-1] >= head[:, 1:])   # sorted per sub-filter up to the received slots
    est = f.estimate()
    assert np.hypot(est.mean[0] - truth[2, 0], est.mean[1] - truth[2, 1]) < 0.5


def test_headline_size_bit_exact_vs_oracle(engine):
    """The 2M-particle headline step itself (64x64 torus of 512, t = 4,
    systematic, 32 landmarks, MTGP32 seed 42; the fused sample+sort launch and
    the two-wave resampling) against the oracle, bit for bit, for two steps."""
    import ctypes
    desc = ea.filter_desc(ea.MODEL_UNICYCLE, m=512, G=4096, t=4, topology=ea.TOPO_TORUS,
                          resampler=ea.RESAMPLE_SYSTEMATIC, weights=ea.WEIGHTS_LOG, seed=42)
    params, ctrl, meas, truth = ea.unicycle_scenario(42, 32, 50.0, 2)
    f = ea.Filter(engine, desc, params)
    assert f.set_fusion(True)
    o = O.Filter(desc, params, ctypes.sizeof(params))
    init = np.zeros(3, np.float32)
    f.init_uniform(init)
    o.init_uniform(init)
    for k in range(2):
        f.step(ctrl[k], meas[k], 0.1)
        o.step(ctrl[k], meas[k], 0.1)
        gs, gw = f.particles()
        os_, ow = o.read(3)
        assert np.array_equal(gs.view(np.uint32), os_.view(np.uint32)), f"state differs at step {k}"
        assert np.array_equal(gw.view(np.uint32), ow.view(np.uint32)), f"weights differ at step {k}"
    est = f.estimate()
    m32, m64, best = o.estimate(3)
    assert est.best_weight == best[0] and est.best_block == int(best[1])
    np.testing.assert_allclose(np.array(est.mean[:3]), m32, rtol=1e-6, atol=1e-6)
    f.close()


def test_maximum_size_properties(engine):
    """The largest filter the descriptor admits: 65536 sub-filters of 1024
    particles (67,108,864 particles, a 256 x 256 torus exchanging 4 per
    neighbour): two steps run and leave finite particles, weights sorted per
    sub-filter up to the received slots, received slots holding the senders'
    best weights, and a posterior mean near the truth; one more sub-filter or
    a larger one is refused."""
    m, G, t = 1024, 65536, 4
    desc = ea.filter_desc(ea.MODEL_UNICYCLE, m=m, G=G, t=t, topology=ea.TOPO_TORUS, weights=ea.WEIGHTS_LOG,
                          resampler=ea.RESAMPLE_SYSTEMATIC, seed=3)
    params, ctrl, meas, truth = ea.unicycle_scenario(3, 8, 20.0, 2)
    f = ea.Filter(engine, desc, params)
    f.init_uniform(np.zeros(3, np.float32))
    for k in range(2):
        f.step(ctrl[k], meas[k], 0.1)
    est = f.estimate()
    st, w = f.particles()
    assert np.isfinite(st).all()
    w = w.reshape(G, m)
    head = w[:, : m - 4 * t]
    assert np.all(head[:, :-1] >= head[:, 1:])
    # every received weight is one of the best t of some sub-filter
    best = np.sort(head[:, :t].reshape(-1))
    recv = w[:, m - 4 * t:].reshape(-1)
    assert np.isin(recv, best).all()
    assert np.hypot(est.mean[0] - truth[1, 0], est.mean[1] - truth[1, 1]) < 0.5
    f.close()
    for bad in (dict(m=2048, G=16), dict(m=1024, G=65537)):
        with pytest.raises(ea.EsthError):
            ea.Filter(engine, ea.filter_desc(ea.MODEL_UNICYCLE, **bad), params)


@pytest.mark.parametrize("graph", [False, True])
def test_prefetch_on_off_identical(graph, engine):
    """Generating step n+1's MTGP words on a second stream during step n
    changes nothing in the results (eager and graph replay)."""
    desc = ea.filter_desc(ea.MODEL_UNICYCLE, m=256, G=64, t=2, topology=ea.TOPO_TORUS, weights=ea.WEIGHTS_LOG,
                          resampler=ea.RESAMPLE_SYSTEMATIC, seed=5)
    params, ctrl, meas, init, truth, dt = make_case(ea.MODEL_UNICYCLE, 9)
    a, b = ea.Filter(engine, desc, params), ea.Filter(engine, desc, params)
    a.set_prefetch(False)
    for f in (a, b):
        f.init_uniform(init)
    if graph:
        b.use_graph(True)
    for k in range(9):
        a.step(ctrl[k], meas[k], dt)
        b.step(ctrl[k], meas[k], dt)
    sa, wa = a.particles()
    sb, wb = b.particles()
    assert np.array_equal(sa.view(np.uint32), sb.view(np.uint32))
    assert np.array_equal(wa.view(np.uint32), wb.view(np.uint32))


FUSION_CASES = [
    ("unicycle-512-soa-torus-log", ea.MODEL_UNICYCLE, 512, ea.LAYOUT_SOA,
     dict(G=16, t=4, topology=ea.TOPO_TORUS, resampler=ea.RESAMPLE_SYSTEMATIC, weights=ea.WEIGHTS_LOG)),
    ("unicycle-128-aos-ring-gate", ea.MODEL_UNICYCLE, 128, ea.LAYOUT_AOS, dict(G=24, t=1, threshold=0.5)),
    ("unicycle-256-packed-carry", ea.MODEL_UNICYCLE, 256, ea.LAYOUT_PACKED,
     dict(G=16, t=2, topology=ea.TOPO_TORUS, threshold=0.5, carry=1, weights=ea.WEIGHTS_LOG)),
    ("unicycle-1024-soa-alias", ea.MODEL_UNICYCLE, 1024, ea.LAYOUT_SOA,
     dict(G=8, t=4, topology=ea.TOPO_TORUS, resampler=ea.RESAMPLE_ALIAS, weights=ea.WEIGHTS_LOG)),
    ("robotarm-128-soa-linear", ea.MODEL_ROBOTARM5, 128, ea.LAYOUT_SOA, dict(G=16, t=1)),
    ("robotarm-256-packed-alias", ea.MODEL_ROBOTARM5, 256, ea.LAYOUT_PACKED,
     dict(G=8, t=1, resampler=ea.RESAMPLE_ALIAS, adaptive=1, threshold=0.01)),
    ("robotarm-512-soa-log", ea.MODEL_ROBOTARM5, 512, ea.LAYOUT_SOA, dict(G=8, t=2, weights=ea.WEIGHTS_LOG)),
    ("dh6-256-soa-torus", ea.MODEL_DH6, 256, ea.LAYOUT_SOA, dict(G=8, t=1, topology=ea.TOPO_TORUS,
                                                                 weights=ea.WEIGHTS_LOG)),
    ("dh6-512-soa-torus", ea.MODEL_DH6, 512, ea.LAYOUT_SOA, dict(G=8, t=1, topology=ea.TOPO_TORUS,
                                                                 weights=ea.WEIGHTS_LOG)),
    ("dh6-1024-soa-ring", ea.MODEL_DH6, 1024, ea.LAYOUT_SOA, dict(G=4, t=1, weights=ea.WEIGHTS_LOG)),
    ("unicycle10-256-soa-all2all", ea.MODEL_UNICYCLE10, 256, ea.LAYOUT_SOA,
     dict(G=16, t=2, topology=ea.TOPO_ALL2ALL, weights=ea.WEIGHTS_LOG)),
    ("camera-256-aos-log", ea.MODEL_ROBOTARM5_CAMERA, 256, ea.LAYOUT_AOS, dict(G=8, t=1, weights=ea.WEIGHTS_LOG)),
]


@pytest.mark.parametrize("name,model,m,layout,kw", FUSION_CASES, ids=[c[0] for c in FUSION_CASES])
def test_fused_sample_sort_identical(name, model, m, layout, kw, engine):
    """Sample+weight and sort in one launch (state kept in LDS) against the
    two launches, step by step: particles, sorted weights and estimates bit
    for bit.  Where the fused launch would cost occupancy it is not taken and
    both filters run the same launches."""
    desc = ea.filter_desc(model, m=m, layout=layout, seed=77, **kw)
    params, ctrl, meas, init, truth, dt = make_case(model, 8)
    a, b = ea.Filter(engine, desc, params), ea.Filter(engine, desc, params)
    fused = a.set_fusion(True)
    assert b.set_fusion(False) is False
    if model in (ea.MODEL_UNICYCLE, ea.MODEL_DH6, ea.MODEL_ROBOTARM5) and m <= 512:
        assert fused, "the fused launch is expected at m <= 512 for these models"
    print(name, "fused" if fused else "not fused (occupancy)", flush=True)
    for f in (a, b):
        f.init_uniform(init)
    for k in range(8):
        d = float(dt[k]) if np.ndim(dt) else dt
        a.step(ctrl[k], meas[k], d)
        b.step(ctrl[k], meas[k], d)
        sa, wa = a.particles()
        sb, wb = b.particles()
        assert np.array_equal(sa.view(np.uint32), sb.view(np.uint32)), f"state differs at step {k}"
        assert np.array_equal(wa.view(np.uint32), wb.view(np.uint32)), f"weights differ at step {k}"
        ea_, eb = a.estimate(), b.estimate()
        assert ea_.best_weight == eb.best_weight and ea_.best_block == eb.best_block
        assert np.array_equal(np.array(ea_.mean), np.array(eb.mean))


def test_fused_sample_sort_graph(engine):
    """The fused launch inside a replayed graph equals eager two-launch steps."""
    desc = ea.filter_desc(ea.MODEL_UNICYCLE, m=512, G=32, t=4, topology=ea.TOPO_TORUS, weights=ea.WEIGHTS_LOG,
                          resampler=ea.RESAMPLE_SYSTEMATIC, seed=9)
    params, ctrl, meas, init, truth, dt = make_case(ea.MODEL_UNICYCLE, 10)
    a, b = ea.Filter(engine, desc, params), ea.Filter(engine, desc, params)
    assert a.set_fusion(False) is False and b.set_fusion(True) is True
    for f in (a, b):
        f.init_uniform(init)
    b.use_graph(True)
    for k in range(10):
        a.step(ctrl[k], meas[k], dt)
        b.step(ctrl[k], meas[k], dt)
        if k % 3 == 2 or k == 9:
            sa, wa = a.particles()
            sb, wb = b.particles()
            assert np.array_equal(sa.view(np.uint32), sb.view(np.uint32))
            assert np.array_equal(wa.view(np.uint32), wb.view(np.uint32))


def test_small_unicycle_500_steps_rmse(engine):
    """The small unicycle workload end to end: 16 x 256 particles, 4 landmarks
    at (+-10, +-10), 4 x 4 torus with one particle per neighbour, systematic
    resampling every step, seed 1234, 500 steps against the seeded ground
    truth.  Particles stay bit-identical to the oracle, posterior means within
    1e-4 of the FP64 twin at every step, and the position RMSE of the
    posterior mean is the oracle's and small."""
    import ctypes
    d = ea.filter_desc(ea.MODEL_UNICYCLE, m=256, G=16, t=1, topology=ea.TOPO_TORUS,
                       resampler=ea.RESAMPLE_SYSTEMATIC, weights=ea.WEIGHTS_LOG, seed=1234)
    p, ctrl, meas, truth = ea.unicycle_scenario(1234, 4, 0.0, 500)
    f = ea.Filter(engine, d, p)
    o = O.Filter(d, p, ctypes.sizeof(p))
    f.init_uniform(np.zeros(3, np.float32))
    o.init_uniform(np.zeros(3, np.float32))
    eg, eo, worst = [], [], 0.0
    for k in range(500):
        f.step(ctrl[k], meas[k], 0.1)
        o.step(ctrl[k], meas[k], 0.1)
        est = f.estimate()
        m32, m64, best = o.estimate(3)
        mean = np.array(est.mean[:3])
        eg.append(np.hypot(mean[0] - truth[k, 0], mean[1] - truth[k, 1]))
        eo.append(np.hypot(m32[0] - truth[k, 0], m32[1] - truth[k, 1]))
        worst = max(worst, float(np.max(np.abs(mean - m64) / (1e-4 * np.abs(m64) + 1e-6))))
    gs, gw = f.particles()
    os_, ow = o.read(3)
    assert np.array_equal(gs.view(np.uint32), os_.view(np.uint32))
    rmse_g, rmse_o = float(np.sqrt(np.mean(np.square(eg)))), float(np.sqrt(np.mean(np.square(eo))))
    assert worst < 1.0
    assert abs(rmse_g - rmse_o) < 1e-6 * max(rmse_o, 1.0)
    assert rmse_g < 0.2, rmse_g
    f.close()


def test_rng_inline_and_stream_paths_identical(engine):
    """The next step's MTGP32 words from extra work-groups of the sampling
    launch (taken where the sampling hides them: the unicycle with 40
    landmarks at m = 512), generated on the main stream (prefetch off), and
    replayed in a graph: the same particles as the oracle at every step,
    across a snapshot taken and restored mid-run; the same filter with 4
    landmarks takes its words from the second stream."""
    import ctypes
    desc = ea.filter_desc(ea.MODEL_UNICYCLE, m=512, G=64, t=4, topology=ea.TOPO_TORUS,
                          resampler=ea.RESAMPLE_SYSTEMATIC, weights=ea.WEIGHTS_LOG, seed=31)
    light, _, _, _ = ea.unicycle_scenario(31, 4, 0.0, 1)
    lf = ea.Filter(engine, desc, light)
    assert lf.rng_mode() == (False, 2)
    lf.close()
    params, ctrl, meas, truth = ea.unicycle_scenario(31, 40, 50.0, 12)
    fs = [ea.Filter(engine, desc, params) for _ in range(3)]
    assert fs[0].rng_mode() == (True, 1), "inline words expected for 40 landmarks at m = 512"
    fs[1].set_prefetch(False)
    fs[2].use_graph(True)
    o = O.Filter(desc, params, ctypes.sizeof(params))
    init = np.zeros(3, np.float32)
    for f in fs:
        f.init_uniform(init)
    o.init_uniform(init)
    snap = None
    for k in range(12):
        for f in fs:
            f.step(ctrl[k], meas[k], 0.1)
        o.step(ctrl[k], meas[k], 0.1)
        os_, ow = o.read(3)
        for i, f in enumerate(fs):
            gs, gw = f.particles()
            assert np.array_equal(gs.view(np.uint32), os_.view(np.uint32)), f"filter {i} state, step {k}"
            assert np.array_equal(gw.view(np.uint32), ow.view(np.uint32)), f"filter {i} weights, step {k}"
        if k == 4:
            snap = fs[0].snapshot()
    r = ea.Filter(engine, desc, params)
    r.restore(snap)
    for k in range(5, 12):
        r.step(ctrl[k], meas[k], 0.1)
    gs, _ = r.particles()
    assert np.array_equal(gs.view(np.uint32), o.read(3)[0].view(np.uint32))
    for f in fs + [r]:
        f.close()
