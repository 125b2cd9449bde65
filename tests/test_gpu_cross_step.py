"""One launch across the step boundary (esth_filter_set_cross_step): step k's
resampling runs at the start of step k + 1's launch.  Whatever the pattern of
reads between steps (none, every step, every few steps, a snapshot, a switch
to graph replay), the particles, weights and estimates equal the oracle's and
the two-launch step's, bit for bit.  The configurations are the unicycle at
m = 512 with 32 landmarks and enough sub-filters that the next step's random
words are generated inline (where the cross-step launch applies)."""
import ctypes as C

import numpy as np
import pytest

import esthera_amd as ea
import oracle as O

pytestmark = pytest.mark.gpu


def same_bits(a, b):
    return np.array_equal(np.asarray(a, np.float32).view(np.uint32), np.asarray(b, np.float32).view(np.uint32))


def make(engine, G, topology, resampler, seed=5, cross=True):
    desc = ea.filter_desc(ea.MODEL_UNICYCLE, m=512, G=G, t=4 if topology == ea.TOPO_TORUS else 2, topology=topology,
                          resampler=resampler, weights=ea.WEIGHTS_LOG, seed=seed)
    p, ctrl, meas, truth = ea.unicycle_scenario(seed, 32, 50.0, 12)
    f = ea.Filter(engine, desc, p)
    f.init_uniform(np.zeros(3, np.float32))
    active = f.set_cross_step(cross)
    return desc, p, ctrl, meas, f, active


def oracle_filter(desc, p):
    o = O.Filter(desc, p, C.sizeof(p))
    o.init_uniform(np.zeros(3, np.float32))
    return o


def check(f, o, k):
    gs, gw = f.particles()
    os_, ow = o.read(3)
    assert same_bits(gs, os_), f"state differs after step {k}"
    assert same_bits(gw, ow), f"weights differ after step {k}"
    est = f.estimate()
    m32, m64, best = o.estimate(3)
    assert est.best_weight == best[0] and est.best_block == int(best[1]), f"best after step {k}"
    assert same_bits(np.array(est.best[:3]), o.best_particle(3))
    np.testing.assert_allclose(np.array(est.mean[:3]), m32, rtol=1e-6, atol=1e-6)


CASES = [("torus-systematic", ea.TOPO_TORUS, ea.RESAMPLE_SYSTEMATIC),
         ("ring-multinomial", ea.TOPO_RING, ea.RESAMPLE_MULTINOMIAL)]


@pytest.mark.parametrize("name,topo,res", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("read_every", [1, 3, 8])
def test_cross_step_equals_oracle(engine, name, topo, res, read_every):
    desc, p, ctrl, meas, f, active = make(engine, 1024, topo, res)
    assert active, "the cross-step launch is expected for this configuration"
    assert f.rng_mode()[0], "next-step words generated inline"
    o = oracle_filter(desc, p)
    steps = 8
    for k in range(steps):
        f.step(ctrl[k], meas[k], 0.1)
        o.step(ctrl[k], meas[k], 0.1)
        if (k + 1) % read_every == 0:
            print(f"check after step {k}", flush=True)
            check(f, o, k)
    f.close()


def test_cross_step_equals_two_launches(engine):
    """The same filter with the cross-step launch on and off, 10 steps with
    no read in between: identical particles and estimates."""
    desc, p, ctrl, meas, a, active = make(engine, 1024, ea.TOPO_TORUS, ea.RESAMPLE_SYSTEMATIC, seed=9)
    _, _, _, _, b, active_b = make(engine, 1024, ea.TOPO_TORUS, ea.RESAMPLE_SYSTEMATIC, seed=9, cross=False)
    assert active and not active_b
    for k in range(10):
        a.step(ctrl[k], meas[k], 0.1)
        b.step(ctrl[k], meas[k], 0.1)
    sa, wa = a.particles()
    sb, wb = b.particles()
    assert same_bits(sa, sb) and same_bits(wa, wb)
    ea_, eb = a.estimate(), b.estimate()
    assert ea_.best_weight == eb.best_weight and ea_.best_block == eb.best_block
    assert np.array_equal(np.array(ea_.mean[:3]), np.array(eb.mean[:3]))
    a.close()
    b.close()


def test_cross_step_snapshot_and_graph_switch(engine):
    """A snapshot taken between cross-step launches and restored later, and a
    switch to graph replay in the middle of a run: the oracle's particles."""
    desc, p, ctrl, meas, f, active = make(engine, 1024, ea.TOPO_TORUS, ea.RESAMPLE_SYSTEMATIC, seed=13)
    assert active
    o = oracle_filter(desc, p)
    for k in range(3):
        f.step(ctrl[k], meas[k], 0.1)
        o.step(ctrl[k], meas[k], 0.1)
    snap = f.snapshot()
    for k in range(3, 5):
        f.step(ctrl[k], meas[k], 0.1)
    f.restore(snap)
    for k in range(3, 6):
        f.step(ctrl[k], meas[k], 0.1)
        o.step(ctrl[k], meas[k], 0.1)
    check(f, o, 5)
    f.use_graph(True)
    for k in range(6, 9):
        f.step(ctrl[k], meas[k], 0.1)
        o.step(ctrl[k], meas[k], 0.1)
    check(f, o, 8)
    f.use_graph(False)
    for k in range(9, 12):
        f.step(ctrl[k], meas[k], 0.1)
        o.step(ctrl[k], meas[k], 0.1)
    check(f, o, 11)
    f.close()
