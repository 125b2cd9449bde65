"""State layouts: a filter whose device-resident state is AOS or PACKED
(the reference's pf_model.cuh / pf_model_soa_packed.cuh variants) steps to
particles, weights and estimates bit-identical to the SOA filter and to the
oracle, for every state dimension, resampler and topology; snapshots carry SOA
planes, so one taken in one layout restores into another; and the layout
autotuner times all three."""
import ctypes as C

import numpy as np
import pytest

import esthera_amd as ea
import oracle as O
from test_gpu_filter import make_case

pytestmark = pytest.mark.gpu

CASES = [
    ("robotarm-multinomial", dict(model=ea.MODEL_ROBOTARM5, m=128, G=16, t=1)),
    ("camera-alias-carry", dict(model=ea.MODEL_ROBOTARM5_CAMERA, m=256, G=8, t=2, resampler=ea.RESAMPLE_ALIAS,
                                weights=ea.WEIGHTS_LOG, threshold=0.6, carry=1)),
    ("unicycle-torus-systematic", dict(model=ea.MODEL_UNICYCLE, m=512, G=16, t=2, topology=ea.TOPO_TORUS,
                                       resampler=ea.RESAMPLE_SYSTEMATIC, weights=ea.WEIGHTS_LOG)),
    ("unicycle-small-m-gate-carry", dict(model=ea.MODEL_UNICYCLE, m=16, G=24, t=2, weights=ea.WEIGHTS_LOG,
                                         threshold=0.5, carry=1)),
    ("dh6-all2all-m1024", dict(model=ea.MODEL_DH6, m=1024, G=4, t=2, topology=ea.TOPO_ALL2ALL,
                               weights=ea.WEIGHTS_LOG)),
    ("unicycle10-ring", dict(model=ea.MODEL_UNICYCLE10, m=64, G=32, t=1)),
]


@pytest.mark.parametrize("layout", [ea.LAYOUT_AOS, ea.LAYOUT_PACKED])
@pytest.mark.parametrize("name,kw", CASES, ids=[c[0] for c in CASES])
def test_layout_steps_bit_identical(name, kw, layout, engine):
    steps = 6
    params, ctrl, meas, init, truth, dt = make_case(kw["model"], steps)
    S = ea.STATE_DIM[kw["model"]]
    ref = ea.Filter(engine, ea.filter_desc(seed=77, **kw), params)
    lay = ea.Filter(engine, ea.filter_desc(seed=77, layout=layout, **kw), params)
    d = ea.filter_desc(seed=77, **kw)
    o = O.Filter(d, params, 0 if params is None else C.sizeof(params))
    for f in (ref, lay, o):
        f.init_uniform(init)
    for k in range(steps):
        dk = float(dt[k]) if np.ndim(dt) else dt
        for f in (ref, lay, o):
            f.step(ctrl[k], meas[k], dk)
        sr, wr = ref.particles()
        sl, wl = lay.particles()
        assert np.array_equal(sl.view(np.uint32), sr.view(np.uint32)), f"state differs at step {k}"
        assert np.array_equal(wl.view(np.uint32), wr.view(np.uint32)), f"weights differ at step {k}"
        er, el = ref.estimate(), lay.estimate()
        assert np.array_equal(np.array(el.mean[:S]), np.array(er.mean[:S]))
        assert el.best_weight == er.best_weight and el.best_block == er.best_block
        assert np.array_equal(np.array(el.best[:S], np.float32), np.array(er.best[:S], np.float32))
    so, wo = o.read(S)
    assert np.array_equal(sl.view(np.uint32), so.view(np.uint32))
    ref.close()
    lay.close()


def test_layout_init_read_roundtrip_and_snapshot(engine):
    model = ea.MODEL_ROBOTARM5
    params, ctrl, meas, init, truth, dt = make_case(model, 8)
    rng = np.random.default_rng(3)
    kw = dict(m=64, G=8, t=1, weights=ea.WEIGHTS_LOG, seed=5)
    N, S = 64 * 8, 9
    soa = rng.normal(size=S * N).astype(np.float32)
    fs = {L: ea.Filter(engine, ea.filter_desc(model, layout=L, **kw), params)
          for L in (ea.LAYOUT_SOA, ea.LAYOUT_AOS, ea.LAYOUT_PACKED)}
    for f in fs.values():
        f.init(soa)
        got, _ = f.particles()
        assert np.array_equal(got, soa.reshape(got.shape))
        for k in range(3):
            f.step(ctrl[k], meas[k], float(dt[k]))
    snap = fs[ea.LAYOUT_AOS].snapshot()
    b = ea.Filter(engine, ea.filter_desc(model, layout=ea.LAYOUT_PACKED, **kw), params)
    b.restore(snap)
    for k in range(3, 8):
        for f in (fs[ea.LAYOUT_SOA], b):
            f.step(ctrl[k], meas[k], float(dt[k]))
    sa, wa = fs[ea.LAYOUT_SOA].particles()
    sb, wb = b.particles()
    assert np.array_equal(sa.view(np.uint32), sb.view(np.uint32))
    assert np.array_equal(wa.view(np.uint32), wb.view(np.uint32))


def test_layout_rejected_by_batches(engine):
    params, ctrl, meas, truth = ea.unicycle_scenario(1, 4, 10.0, 2)
    d = ea.filter_desc(ea.MODEL_UNICYCLE, m=64, G=2, rng=ea.RNG_MTGP_PER_FILTER, layout=ea.LAYOUT_AOS)
    with pytest.raises(ea.EsthError):
        ea.Batch(engine, [d], [params])
    with pytest.raises(ea.EsthError):
        ea.Filter(engine, ea.filter_desc(ea.MODEL_UNICYCLE, layout=3), params)


def test_autotune_layout_times_all_three(engine):
    params, ctrl, meas, init, truth, dt = make_case(ea.MODEL_ROBOTARM5, 2)
    d = ea.filter_desc(ea.MODEL_ROBOTARM5, m=256, G=256, t=1)
    us, best = ea.autotune_layout(engine, d, params, ctrl[0], meas[0], float(dt[0]))
    assert us.shape == (3,) and np.all(us > 0) and np.all(np.isfinite(us))
    assert best == int(np.argmin(us))
