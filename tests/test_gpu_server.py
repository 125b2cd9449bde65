"""The served form: tracks of different models, sizes and settings opened,
updated and closed by requests that arrive while earlier waves run.  Every
result (the estimate of each update) and every track's final particles must be
bit-identical to the same filter stepped alone with per-request launches, and
to the oracle's run of that filter (posterior means within
1e-4 * |m64| + 1e-6 of its FP64 twin)."""
import ctypes
import os

import numpy as np
import pytest

import esthera_amd as ea
import oracle as O

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
GiB = 1 << 30


def unicycle_track(seed, m, G, L=8, topology=ea.TOPO_TORUS, t=1, resampler=ea.RESAMPLE_SYSTEMATIC, threshold=1.0,
                   carry=0):
    d = ea.filter_desc(ea.MODEL_UNICYCLE, m=m, G=G, t=t, topology=topology, resampler=resampler,
                       weights=ea.WEIGHTS_LOG, rng=ea.RNG_MTGP_PER_FILTER, seed=seed, threshold=threshold,
                       carry=carry)
    p, ctrl, meas, truth = ea.unicycle_scenario(seed, L, 20.0, 40)
    return d, p, [(ctrl[k], meas[k], 0.1) for k in range(40)], np.zeros(3, np.float32)


def robotarm_track(seed, m, G, resampler=ea.RESAMPLE_MULTINOMIAL):
    d = ea.filter_desc(ea.MODEL_ROBOTARM5, m=m, G=G, t=1, topology=ea.TOPO_RING, resampler=resampler,
                       weights=ea.WEIGHTS_LOG, rng=ea.RNG_MTGP_PER_FILTER, seed=seed)
    meas, ctrl, dts, truth = ea.load_robotarm_trace(os.path.join(HERE, "golden", "robotarm_input5.txt"))
    init = np.array([0.75 * np.pi] * 5 + [7.0, 0.0, 0.0, 0.0], np.float32)
    return d, None, [(ctrl[k], meas[k], float(dts[k])) for k in range(40)], init


def alone(engine, d, p, inputs, init, n):
    f = ea.Filter(engine, d, p)
    f.init_uniform(init)
    ests = []
    for k in range(n):
        f.step(*inputs[k])
        ests.append(f.estimate())
    st, w = f.particles()
    f.close()
    return ests, st, w


def same_estimate(a, b, S):
    """Best particle exact; the posterior sums are reduced over sub-filters in
    double by 256 threads here and 1024 in the single-filter kernel, so the
    means agree to double rounding."""
    return (np.allclose(np.array(a.mean[:S]), np.array(b.mean[:S]), rtol=1e-12, atol=1e-12) and
            a.best_weight == b.best_weight and a.best_block == b.best_block and
            np.isclose(a.weight_sum, b.weight_sum, rtol=1e-12, atol=0) and
            np.array_equal(np.array(a.best[:S]), np.array(b.best[:S])))


def test_server_mix_equals_tracks_alone(engine):
    specs = [unicycle_track(11, 256, 16), unicycle_track(12, 64, 40), robotarm_track(13, 128, 8),
             unicycle_track(14, 1024, 4, L=32), unicycle_track(15, 512, 6, topology=ea.TOPO_RING, t=2,
                                                               resampler=ea.RESAMPLE_ALIAS),
             robotarm_track(16, 256, 3), unicycle_track(17, 128, 9, topology=ea.TOPO_ALL2ALL, t=2),
             unicycle_track(18, 256, 6, threshold=0.5, carry=1), unicycle_track(19, 128, 8, threshold=0.5)]
    nsteps = [9, 14, 7, 11, 6, 12, 8, 10, 9]
    srv = ea.Server(engine, 64, GiB)
    tracks = [None] * len(specs)
    sent = [0] * len(specs)
    expect = {}
    rng = np.random.default_rng(5)
    # staggered: tracks open over time, updates arrive in random order while waves run
    for rnd in range(40):
        for i, (d, p, inputs, init) in enumerate(specs):
            if tracks[i] is None and rnd >= i:
                tracks[i], _ = srv.open(d, p, init)
            elif tracks[i] is not None and sent[i] < nsteps[i] and rng.random() < 0.6:
                rid = srv.update(tracks[i], *inputs[sent[i]])
                expect[rid] = (i, sent[i])
                sent[i] += 1
        srv.pump()
    for i, (d, p, inputs, init) in enumerate(specs):
        while sent[i] < nsteps[i]:
            rid = srv.update(tracks[i], *inputs[sent[i]])
            expect[rid] = (i, sent[i])
            sent[i] += 1
    srv.drain()
    results = srv.poll(20000)
    got = {}
    for r in results:
        assert r.status == 0, (r.kind, r.status)
        if r.kind == 1:
            got[r.request_id] = r
    assert set(got) == set(expect)
    ref = [alone(engine, d, p, inputs, init, nsteps[i]) for i, (d, p, inputs, init) in enumerate(specs)]
    for rid, (i, k) in expect.items():
        S = ea.STATE_DIM[specs[i][0].model]
        assert got[rid].estimate.step == k + 1
        assert same_estimate(got[rid].estimate, ref[i][0][k], S), f"track {i} step {k}"
    for i in range(len(specs)):
        st, w = srv.particles(tracks[i])
        assert np.array_equal(st.view(np.uint32), ref[i][1].view(np.uint32)), f"track {i} state"
        assert np.array_equal(w.view(np.uint32), ref[i][2].view(np.uint32)), f"track {i} weights"
    by_step = {v: rid for rid, v in expect.items()}
    for i, (d, p, inputs, init) in enumerate(specs):
        S = ea.STATE_DIM[d.model]
        o = O.Filter(d, p, ctypes.sizeof(p) if p is not None else 0)
        o.init_uniform(init)
        for k in range(nsteps[i]):
            o.step(*inputs[k])
            m32, m64, best = o.estimate(S)
            e = got[by_step[(i, k)]].estimate
            assert e.best_weight == best[0] and e.best_block == int(best[1]), f"track {i} step {k}"
            mean = np.array(e.mean[:S])
            assert np.all(np.abs(mean - m64) <= 1e-4 * np.abs(m64) + 1e-6), f"track {i} step {k}"
        ost, ow = o.read(S)
        st, w = srv.particles(tracks[i])
        assert np.array_equal(st.view(np.uint32), ost.view(np.uint32)), f"track {i} state vs oracle"
        assert np.array_equal(w.view(np.uint32), ow.view(np.uint32)), f"track {i} weights vs oracle"
    waves, wave_updates, peak = srv.wave_stats()
    assert wave_updates == len(expect) and waves < len(expect)
    srv.close()


def test_admission_waits_for_memory_and_reuses_blocks(engine):
    d, p, inputs, init = unicycle_track(21, 256, 64)   # 16k particles
    srv = ea.Server(engine, 64, 2 << 20)                # room for about four such tracks
    ids = [srv.open(ea.filter_desc(ea.MODEL_UNICYCLE, m=256, G=64, t=1, topology=ea.TOPO_TORUS,
                                   resampler=ea.RESAMPLE_SYSTEMATIC, weights=ea.WEIGHTS_LOG,
                                   rng=ea.RNG_MTGP_PER_FILTER, seed=100 + j), p, init)[0] for j in range(12)]
    srv.pump()
    used, live, queued = srv.stats()
    assert 0 < live < 12 and queued > 0
    for tid in ids:
        for k in range(3):
            srv.update(tid, *inputs[k])
        srv.close_track(tid)
    for _ in range(200):
        srv.pump()
        if srv.stats()[2] == 0:
            break
    srv.drain()
    res = srv.poll(20000)
    assert all(r.status == 0 for r in res)
    assert sum(r.kind == 1 for r in res) == 36 and sum(r.kind == 2 for r in res) == 12
    by = {}
    for r in res:
        if r.kind == 1:
            by.setdefault(r.track_id, []).append(r)
    for j, tid in enumerate(ids):
        dj = ea.filter_desc(ea.MODEL_UNICYCLE, m=256, G=64, t=1, topology=ea.TOPO_TORUS,
                            resampler=ea.RESAMPLE_SYSTEMATIC, weights=ea.WEIGHTS_LOG,
                            rng=ea.RNG_MTGP_PER_FILTER, seed=100 + j)
        ests, _, _ = alone(engine, dj, p, inputs, init, 3)
        rs = sorted(by[tid], key=lambda r: r.request_id)
        assert [r.estimate.step for r in rs] == [1, 2, 3]
        for k in range(3):
            assert same_estimate(rs[k].estimate, ests[k], 3), f"track {j} step {k}"
    assert srv.stats()[0] == 0 and srv.stats()[1] == 0
    srv.close()


def test_request_errors(engine):
    d, p, inputs, init = unicycle_track(31, 256, 4)
    srv = ea.Server(engine, 4, 64 << 20)
    with pytest.raises(ea.EsthError):
        srv.update(12345, *inputs[0])
    bad = ea.filter_desc(ea.MODEL_UNICYCLE, m=32, G=4, t=1, weights=ea.WEIGHTS_LOG, rng=ea.RNG_MTGP_PER_FILTER)
    with pytest.raises(ea.EsthError):
        srv.open(bad, p, init)
    compat = ea.filter_desc(ea.MODEL_UNICYCLE, m=256, G=4, t=1, weights=ea.WEIGHTS_LOG, rng=ea.RNG_MTGP_COMPAT)
    with pytest.raises(ea.EsthError):
        srv.open(compat, p, init)
    huge = ea.filter_desc(ea.MODEL_UNICYCLE, m=1024, G=4096, t=1, weights=ea.WEIGHTS_LOG, rng=ea.RNG_MTGP_PER_FILTER)
    with pytest.raises(ea.EsthError):
        srv.open(huge, p, init)
    tid, _ = srv.open(d, p, init)
    srv.update(tid, *inputs[0])
    srv.close_track(tid)
    with pytest.raises(ea.EsthError):
        srv.update(tid, *inputs[1])
    srv.drain()
    kinds = sorted(r.kind for r in srv.poll())
    assert kinds == [0, 1, 2]
    srv.close()


def test_serve_simulate_small(engine):
    meas, ctrl, dts, truth = ea.load_robotarm_trace(os.path.join(HERE, "golden", "robotarm_input5.txt"))
    w = ea.ServeWorkload(rate=2000.0, duration_s=0.5, open_fraction=0.05, unicycle_fraction=0.7,
                         min_particles=1024, max_particles=16384, mean_lifetime=50.0, initial_tracks=16,
                         landmarks=8, topology=ea.TOPO_TORUS, num_transfer=1, resampler=ea.RESAMPLE_SYSTEMATIC,
                         weights=ea.WEIGHTS_LOG, seed=3, plain=0, hbm_bytes=2 * GiB, max_tracks=4096,
                         max_inflight=2)
    r = ea.serve_simulate(engine, w, (meas, ctrl, dts))
    assert r["failed"] == 0 and r["updates"] > 500 and r["backlog"] == 0
    assert 0 < r["p50_us"] <= r["p99_us"] < 1e6
    w.plain = 1
    w.rate = 0.0
    w.duration_s = 0.3
    rp = ea.serve_simulate(engine, w, (meas, ctrl, dts))
    assert rp["updates"] > 0 and rp["failed"] == 0
