"""Stage-by-stage parity of the gfx950 kernels against the CPU oracle, called
through the C ABI on device buffers.  Integer/byte/index outputs (MTGP words,
permutations, selections, alias tables, exchanged slots) and, because both
sides evaluate the same IEEE operation sequence, float outputs are compared
bit for bit."""
import ctypes as C

import numpy as np
import pytest

import esthera_amd as ea
import oracle as O

pytestmark = pytest.mark.gpu

DB = ea.DeviceBuffer


def f32bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


def assert_bits(a, b):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape
    if a.dtype == np.float32:
        a, b = f32bits(a), f32bits(b)
    bad = np.nonzero(a != b)[0]
    assert bad.size == 0, f"{bad.size} mismatches, first at {bad[:5]}: {a[bad[:5]]} vs {b[bad[:5]]}"


# ----------------------------------------------------------------- MTGP32
def test_mtgp_16x256_buffer_matches_reference_library(golden, engine):
    """96 streams x 768 words, seed 1234: the reference's first-step buffer."""
    st = np.concatenate([ea.mtgp_seed_state(i, 1234 + i + 1) for i in range(96)])
    dst = DB.from_array(st)
    out = DB(96 * 768 * 4)
    ea.check(ea.lib().esth_stage_mtgp(dst.ptr, 96, 0, 768, out.ptr, 1))
    got = out.to_array(np.uint32, 96 * 768).reshape(96, 768)
    assert_bits(got.ravel(), golden["buf16x256_step1"].ravel())


def test_mtgp_u32_and_continuation(golden, engine):
    """Tempered uint32 words over several launches equal one long sequential run."""
    st = np.concatenate([ea.mtgp_seed_state(i, 1234 + i + 1) for i in range(4)])
    dst = DB.from_array(st)
    chunks = []
    out = DB(4 * 1536 * 4)
    for _ in range(2):
        ea.check(ea.lib().esth_stage_mtgp(dst.ptr, 4, 0, 1536, out.ptr, 0))
        chunks.append(out.to_array(np.uint32, 4 * 1536).reshape(4, 1536))
    got = np.concatenate(chunks, axis=1)
    assert_bits(got.ravel(), golden["u32_1234"][:, :3072].ravel())


@pytest.mark.parametrize("seed", [42, 7, 99, 0xF000000D])
def test_mtgp_all_200_sets(golden, seed, engine):
    st = np.concatenate([ea.mtgp_seed_state(i, (seed + i + 1) & 0xFFFFFFFF) for i in range(200)])
    dst = DB.from_array(st)
    out = DB(200 * 768 * 4)
    ea.check(ea.lib().esth_stage_mtgp(dst.ptr, 200, 0, 768, out.ptr, 1))
    got = out.to_array(np.uint32, 200 * 768).reshape(200, 768)[:, :16]
    assert_bits(got.ravel(), golden[f"f32_{seed}"].ravel())


@pytest.mark.parametrize("nsteps,size", [(1, 768), (3, 1536), (8, 768 * 5)])
def test_mtgp_steps_equal_consecutive_launches(golden, nsteps, size, engine):
    """A block of steps in one launch (how filters generate their words)
    equals one launch per step: every step's words and the status after every
    step; the first 4096 words of each stream are the reference library's."""
    ns = 6
    st = np.concatenate([ea.mtgp_seed_state(i, 1234 + i + 1) for i in range(ns)])
    din, dst_out = DB.from_array(st), DB(nsteps * ns * 351 * 4)
    out = DB(nsteps * ns * size * 4)
    ea.check(ea.lib().esth_stage_mtgp_steps(din.ptr, dst_out.ptr, ns, 0, size, nsteps, out.ptr, 0))
    words = out.to_array(np.uint32, nsteps * ns * size).reshape(nsteps, ns, size)
    stats = dst_out.to_array(np.uint32, nsteps * ns * 351).reshape(nsteps, ns * 351)
    one = DB.from_array(st)
    o1 = DB(ns * size * 4)
    for k in range(nsteps):
        ea.check(ea.lib().esth_stage_mtgp(one.ptr, ns, 0, size, o1.ptr, 0))
        assert_bits(words[k].ravel(), o1.to_array(np.uint32, ns * size))
        assert_bits(stats[k], one.to_array(np.uint32, ns * 351))
    seq = np.concatenate([words[k] for k in range(nsteps)], axis=1)
    n = min(seq.shape[1], 4096)
    assert_bits(seq[:4, :n].ravel(), golden["u32_1234"][:, :n].ravel())


def test_mtgp_state_written_back(mtgp_table, engine):
    st = np.concatenate([ea.mtgp_seed_state(i, 77 + i) for i in range(3)])
    dst = DB.from_array(st)
    out = DB(3 * 2304 * 4)
    ea.check(ea.lib().esth_stage_mtgp(dst.ptr, 3, 0, 2304, out.ptr, 1))
    gst = dst.to_array(np.uint32, 3 * 351)
    ost = st.copy()
    O.mtgp_fill(mtgp_table, 3, ost, 2304)
    assert_bits(gst, ost)


# ----------------------------------------------------------- Box-Muller, math
def test_boxmuller_bit_exact(engine):
    rng = np.random.default_rng(3)
    u = (1.0 + rng.random(200_002, dtype=np.float32)).astype(np.float32)
    u[:4] = [1.0, np.float32(2.0 - 2 ** -23), 1.5, np.float32(1.0 + 2 ** -23)]
    d = DB.from_array(u)
    ea.check(ea.lib().esth_stage_boxmuller(d.ptr, u.size))
    got = d.to_array(np.float32, u.size)
    assert_bits(got, O.boxmuller(u))
    u1 = 2.0 - u[0::2].astype(np.float64)
    u2 = 2.0 - u[1::2].astype(np.float64)
    r = np.sqrt(-2.0 * np.log(u1))
    ref = np.empty(u.size)
    ref[0::2] = r * np.cos(2 * np.pi * u2)
    ref[1::2] = r * np.sin(2 * np.pi * u2)
    assert np.max(np.abs(got - ref)) < 2e-6


@pytest.mark.parametrize("fn", [0, 1, 2, 3, 4, 5, 6, 7])
def test_math_bit_exact(fn, engine):
    rng = np.random.default_rng(fn)
    n = 1 << 16
    if fn == 0:
        x = rng.uniform(-100, 90, n).astype(np.float32)
    elif fn == 1:
        x = np.exp(rng.uniform(-90, 88, n)).astype(np.float32)
    elif fn in (2, 3):
        x = rng.uniform(-1000, 1000, n).astype(np.float32)
    elif fn == 7:
        x = np.exp(rng.uniform(-80, 80, n)).astype(np.float32)
    else:
        x = rng.normal(size=n).astype(np.float32)
    x2 = rng.normal(size=n).astype(np.float32)
    dx, dx2, do = DB.from_array(x), DB.from_array(x2), DB(n * 4)
    ea.check(ea.lib().esth_stage_math(fn, dx.ptr, dx2.ptr, do.ptr, n))
    got = do.to_array(np.float32, n)
    L = O.lib()
    if fn == 0:
        ref = np.array([L.or_expf(v) for v in x], np.float32)
    elif fn == 1:
        ref = np.array([L.or_logf(v) for v in x], np.float32)
    elif fn == 4:
        ref = np.array([L.or_atan2f(a, b) for a, b in zip(x, x2)], np.float32)
    elif fn in (5, 6):
        import ctypes as C
        r, a = C.c_float(), C.c_float()
        ref = []
        for uu, vv in zip(x, x2):
            L.or_polar_abs(uu, vv, C.byref(r), C.byref(a))
            ref.append(r.value if fn == 5 else a.value)
        ref = np.array(ref, np.float32)
    elif fn == 7:
        ref = np.array([L.or_rcpf(v) for v in x], np.float32)
    else:
        import ctypes as C
        s, c = C.c_float(), C.c_float()
        ref = []
        for v in x:
            L.or_sincosf(v, C.byref(s), C.byref(c))
            ref.append(s.value if fn == 2 else c.value)
        ref = np.array(ref, np.float32)
    assert_bits(got, ref)


# ---------------------------------------------------------- sample + weight
def _model_case(model, N, seed):
    rng = np.random.default_rng(seed)
    S, V = ea.STATE_DIM[model], ea.NOISE_DIM[model]
    if model == ea.MODEL_ROBOTARM5:
        params = None
        ctrl = ea.RobotArm5Control()
        meas = ea.RobotArm5Measurement()
        for i in range(5):
            ctrl.angles[i] = rng.normal() * 0.05
            meas.angles[i] = 2.35 + rng.normal() * 0.05
        meas.x, meas.y = -9.9, 11.1
        st = np.concatenate([np.full(N, 2.356, np.float32) + rng.normal(0, 0.05, N).astype(np.float32)
                             for _ in range(5)] + [rng.normal(7, 0.5, N).astype(np.float32),
                                                   rng.normal(0, 0.5, N).astype(np.float32),
                                                   rng.normal(0, 1, N).astype(np.float32),
                                                   rng.normal(0, 1, N).astype(np.float32)])
        dt = 0.0589
    elif model == ea.MODEL_UNICYCLE:
        params, ctrls, meass, truth = ea.unicycle_scenario(42, 32, 50.0, 3)
        ctrl, meas = ctrls[2], meass[2]
        st = np.concatenate([truth[1, c] + rng.normal(0, 0.3, N).astype(np.float32) for c in range(3)])
        dt = 0.1
    elif model == ea.MODEL_UNICYCLE10:
        import os
        meass, ctrls, dts, truth = ea.load_unicycle10_trace(
            os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "unicycle_input.txt"))
        params, ctrl, meas, dt = ea.unicycle10_params_default(), ctrls[2], meass[2], float(dts[2])
        st = np.concatenate([truth[1, c] + rng.normal(0, 0.3, N).astype(np.float32) for c in range(3)])
    elif model == ea.MODEL_ROBOTARM5_CAMERA:
        params = ea.camera_params_default()
        ctrls, meass, truth = ea.robotarm5_camera_scenario(3, 4, 0.1)
        ctrl, meas, dt = ctrls[2], meass[2], 0.1
        st = np.concatenate([truth[2, c] + rng.normal(0, 0.05, N).astype(np.float32) for c in range(9)])
    elif model == ea.MODEL_RANDOMWALK3:
        params, ctrls, meass, truth = ea.randomwalk3_scenario(4, 3)
        ctrl, meas, dt = ctrls[2], meass[2], 0.1
        st = np.concatenate([truth[1, c] + rng.normal(0, 0.2, N).astype(np.float32) for c in range(3)])
    else:
        params, meass, truth = ea.dh6_scenario(7, 3)
        ctrl, meas = ea.DH6Control(), meass[2]
        st = np.concatenate([truth[2, c] + rng.normal(0, 0.02, N).astype(np.float32) for c in range(12)])
        dt = 0.01
    R = (1.0 + rng.random(N * V, dtype=np.float32)).astype(np.float32)
    return params, ctrl, meas, st.astype(np.float32), R, dt, S, V


@pytest.mark.parametrize("model", [ea.MODEL_ROBOTARM5, ea.MODEL_UNICYCLE, ea.MODEL_DH6, ea.MODEL_UNICYCLE10,
                                   ea.MODEL_ROBOTARM5_CAMERA, ea.MODEL_RANDOMWALK3])
@pytest.mark.parametrize("weights", [ea.WEIGHTS_LINEAR, ea.WEIGHTS_LOG])
def test_sample_weight_bit_exact(model, weights, engine):
    N = 4096
    params, ctrl, meas, st, R, dt, S, V = _model_case(model, N, 11 + model)
    fr = ea.build_frame(model, params, ctrl, meas, dt)
    assert_bits(fr, O.build_frame(model, params, ctrl, meas, dt))
    dst, dR, dw = DB.from_array(st), DB.from_array(R), DB(N * 4)
    ea.check(ea.lib().esth_stage_sample_weight(model, dst.ptr, N, dR.ptr, ea.ptr(fr), weights, dw.ptr))
    gst, gw = dst.to_array(np.float32, S * N), dw.to_array(np.float32, N)
    normals = O.boxmuller(R)
    ost, ow = O.sample_weight(model, st, N, normals, fr, weights)
    assert_bits(gst, ost)
    assert_bits(gw, ow)
    # FP64 twin: the model's definition evaluated in double
    s64, lw64 = O.sample_weight_f64(model, st, N, normals, params, ctrl, meas, dt, S)
    np.testing.assert_allclose(gst, s64, rtol=1e-5, atol=2e-5)
    lw = gw if weights == ea.WEIGHTS_LOG else np.log(np.maximum(gw, 1e-38))
    ok = lw64 > -80
    np.testing.assert_allclose(lw[ok], lw64[ok], rtol=2e-4, atol=2e-3)


# ---------------------------------------------------------------- sort
SORT_CASES = [(2, 5), (4, 3), (8, 9), (16, 4), (32, 3), (64, 5), (128, 3), (256, 16), (512, 7), (1024, 3)]


@pytest.mark.parametrize("m,G", SORT_CASES)
@pytest.mark.parametrize("S", [1, 3, 9, 12])
def test_sort_bit_exact_with_ties(m, G, S, engine):
    rng = np.random.default_rng(m * 31 + G + S)
    N = m * G
    w = rng.integers(0, 5, N).astype(np.float32) * 0.25   # many ties
    w[: N // 4] = rng.random(N // 4, dtype=np.float32)
    if N > 8:
        w[3] = 0.0
        w[5] = -0.0
    st = rng.normal(size=S * N).astype(np.float32)
    dw, dst, dout = DB.from_array(w), DB.from_array(st), DB(S * N * 4)
    dwin, dperm = DB(G * 8), DB(N * 4)
    ea.check(ea.lib().esth_stage_sort(dw.ptr, dst.ptr, dout.ptr, S, m, G, dwin.ptr, dperm.ptr))
    ow, oout, owin, operm = O.block_sort(w, st, S, m, G)
    assert_bits(dw.to_array(np.float32, N), ow)
    assert_bits(dperm.to_array(np.int32, N), operm)
    assert_bits(dout.to_array(np.float32, S * N), oout)
    assert_bits(dwin.to_array(np.float32, 2 * G), owin)
    sw = ow.reshape(G, m)
    assert np.all(sw[:, :-1] >= sw[:, 1:])


def test_sort_brute_force_all_tie_patterns(engine):
    """m = 8: every key pattern over {0,1,2}^8 (6561 sub-filters in one launch)."""
    m = 8
    pats = np.array(np.meshgrid(*[np.arange(3)] * m, indexing="ij")).reshape(m, -1).T.astype(np.float32)
    G = pats.shape[0]
    w = pats.ravel().copy()
    dw, dperm, dst, dout = DB.from_array(w), DB(w.size * 4), DB.from_array(np.zeros(w.size, np.float32)), DB(w.size * 4)
    ea.check(ea.lib().esth_stage_sort(dw.ptr, dst.ptr, dout.ptr, 1, m, G, None, dperm.ptr))
    ow, _, _, operm = O.block_sort(w, np.zeros(w.size, np.float32), 1, m, G)
    assert_bits(dperm.to_array(np.int32, w.size), operm)
    gw = dw.to_array(np.float32, w.size).reshape(G, m)
    assert np.array_equal(gw, -np.sort(-pats, axis=1))


# --------------------------------------------------------------- argmax
@pytest.mark.parametrize("G", [1, 16, 127, 128, 129, 4096, 65536])
@pytest.mark.parametrize("weights", [ea.WEIGHTS_LINEAR, ea.WEIGHTS_LOG])
def test_argmax_reference_tie_rule(G, weights, engine):
    rng = np.random.default_rng(G)
    win = np.zeros(2 * G, np.float32)
    win[0::2] = rng.integers(0, 4, G).astype(np.float32) - (5.0 if weights == ea.WEIGHTS_LOG else 0.0)
    win[1::2] = np.arange(G, dtype=np.float32)
    dwin, dres = DB.from_array(win), DB(8)
    ea.check(ea.lib().esth_stage_argmax(dwin.ptr, G, weights, dres.ptr))
    got = dres.to_array(np.float32, 2)
    assert_bits(got, O.argmax(win, G, weights))
    assert got[0] == win[0::2].max()


# ------------------------------------------------------------- exchange
@pytest.mark.parametrize("topo,m,G,t", [(ea.TOPO_RING, 256, 16, 1), (ea.TOPO_RING, 64, 7, 32),
                                        (ea.TOPO_TORUS, 512, 4096, 4), (ea.TOPO_TORUS, 512, 2048, 1),
                                        (ea.TOPO_TORUS, 16, 16, 4), (ea.TOPO_TORUS, 64, 2, 2),
                                        (ea.TOPO_ALL2ALL, 256, 100, 8), (ea.TOPO_ALL2ALL, 128, 4096, 4)])
def test_exchange_bit_exact(topo, m, G, t, engine):
    rng = np.random.default_rng(m + G + t)
    S, N = 3, m * G
    w = -np.sort(-rng.integers(0, 50, N).astype(np.float32).reshape(G, m), axis=1).ravel()
    st = rng.normal(size=S * N).astype(np.float32)
    dw, dst = DB.from_array(w), DB.from_array(st)
    ea.check(ea.lib().esth_stage_exchange(dst.ptr, dw.ptr, S, m, G, t, topo, 0, 0))
    ost, ow = O.exchange(st, w, S, m, G, t, topo)
    assert_bits(dw.to_array(np.float32, N), ow)
    assert_bits(dst.to_array(np.float32, S * N), ost)


# ------------------------------------------------------------ resampling
@pytest.mark.parametrize("resampler", [ea.RESAMPLE_MULTINOMIAL, ea.RESAMPLE_SYSTEMATIC, ea.RESAMPLE_ALIAS])
@pytest.mark.parametrize("weights", [ea.WEIGHTS_LINEAR, ea.WEIGHTS_LOG])
@pytest.mark.parametrize("m,G", [(2, 33), (16, 9), (64, 5), (256, 16), (512, 64), (1024, 3)])
def test_resample_bit_exact(resampler, weights, m, G, engine):
    rng = np.random.default_rng(m + G + 7 * resampler + weights)
    S, N = 3, m * G
    if weights == ea.WEIGHTS_LOG:
        w = (-rng.exponential(20.0, N)).astype(np.float32)
    else:
        w = rng.random(N, dtype=np.float32) ** 4
        w[: m] = 0.0                        # an all-zero sub-filter
    src = rng.normal(size=S * N).astype(np.float32)
    dst0 = rng.normal(size=S * N).astype(np.float32)
    nr = (2 * m + 1) * G if resampler == ea.RESAMPLE_ALIAS else (m + 1) * G
    rnd = (1.0 + rng.random(nr, dtype=np.float32)).astype(np.float32)
    thr = 0.7
    d_dst, d_src, d_w, d_r, d_sel = DB.from_array(dst0), DB.from_array(src), DB.from_array(w), DB.from_array(rnd), DB(N * 4)
    ea.check(ea.lib().esth_stage_resample(resampler, d_dst.ptr, d_src.ptr, S, d_w.ptr, d_r.ptr, m, G, thr, 0,
                                          weights, d_sel.ptr))
    od, osel = O.resample(resampler, dst0, src, S, w, rnd, m, G, thr, 0, weights)
    assert_bits(d_sel.to_array(np.int32, N), osel)
    assert_bits(d_dst.to_array(np.float32, S * N), od)


@pytest.mark.parametrize("thr", [0.0, 0.01, 0.2])
def test_alias_adaptive_bit_exact(thr, engine):
    rng = np.random.default_rng(5)
    m, G, S = 256, 32, 3
    N = m * G
    w = rng.random(N, dtype=np.float32) ** 8
    src = rng.normal(size=S * N).astype(np.float32)
    dst0 = rng.normal(size=S * N).astype(np.float32)
    rnd = (1.0 + rng.random(2 * m * G, dtype=np.float32)).astype(np.float32)
    d_dst, d_src, d_w, d_r, d_sel = DB.from_array(dst0), DB.from_array(src), DB.from_array(w), DB.from_array(rnd), DB(N * 4)
    ea.check(ea.lib().esth_stage_resample(ea.RESAMPLE_ALIAS, d_dst.ptr, d_src.ptr, S, d_w.ptr, d_r.ptr, m, G, thr, 1,
                                          ea.WEIGHTS_LINEAR, d_sel.ptr))
    od, osel = O.resample(ea.RESAMPLE_ALIAS, dst0, src, S, w, rnd, m, G, thr, 1, ea.WEIGHTS_LINEAR)
    assert_bits(d_sel.to_array(np.int32, N), osel)
    assert_bits(d_dst.to_array(np.float32, S * N), od)


@pytest.mark.parametrize("m,G", [(2, 10), (64, 8), (512, 16), (1024, 4)])
def test_alias_table_bit_exact_and_invariant(m, G, engine):
    rng = np.random.default_rng(m)
    w = (rng.random(m * G, dtype=np.float32) ** 3).astype(np.float32)
    dw, dp, da = DB.from_array(w), DB(m * G * 4), DB(m * G * 4)
    ea.check(ea.lib().esth_stage_alias_table(dw.ptr, m, G, ea.WEIGHTS_LINEAR, dp.ptr, da.ptr))
    prob, alias = dp.to_array(np.float32, m * G), da.to_array(np.uint32, m * G)
    op, oa = O.alias_table(w, m, G, ea.WEIGHTS_LINEAR)
    assert_bits(prob, op)
    assert_bits(alias, oa)
    for g in range(G):
        p = w[g * m:(g + 1) * m].astype(np.float64)
        p /= p.sum()
        pr = np.clip(prob[g * m:(g + 1) * m].astype(np.float64), 0, 1)
        mass = pr.copy()
        np.add.at(mass, alias[g * m:(g + 1) * m], 1.0 - pr)
        np.testing.assert_allclose(mass, m * p, atol=2e-5 * m)


@pytest.mark.parametrize("shape", ["peaked", "two-heavy", "flat", "dead"])
@pytest.mark.parametrize("m", [32, 64, 128, 1024])
def test_alias_table_pairing_regimes(shape, m, engine):
    """Weights whose pairing runs many short rounds (one or two heavy
    particles), equal weights (no round at all) and all weights 0 (log -inf:
    the uniform table), in log mode; m = 32 runs the work-group kernel, the
    others one wave per sub-filter."""
    rng = np.random.default_rng(m + len(shape))
    G = 6
    if shape == "peaked":
        lw = (30.0 * rng.random(m * G)).astype(np.float32)
    elif shape == "two-heavy":
        lw = rng.normal(size=m * G).astype(np.float32)
        lw[::m] += 9.0
        lw[m // 2::m] += 8.0
    elif shape == "flat":
        lw = np.full(m * G, -3.25, np.float32)
    else:
        lw = np.full(m * G, -np.inf, np.float32)
    dw, dp, da = DB.from_array(lw), DB(m * G * 4), DB(m * G * 4)
    ea.check(ea.lib().esth_stage_alias_table(dw.ptr, m, G, ea.WEIGHTS_LOG, dp.ptr, da.ptr))
    prob, alias = dp.to_array(np.float32, m * G), da.to_array(np.uint32, m * G)
    op, oa = O.alias_table(lw, m, G, ea.WEIGHTS_LOG)
    assert_bits(prob, op)
    assert_bits(alias, oa)


# ------------------------------------------------- arithmetic floor timing
@pytest.mark.parametrize("model", [ea.MODEL_ROBOTARM5, ea.MODEL_UNICYCLE, ea.MODEL_DH6, ea.MODEL_UNICYCLE10,
                                   ea.MODEL_ROBOTARM5_CAMERA, ea.MODEL_RANDOMWALK3])
def test_time_hook(model, engine):
    """esth_engine_time_hook: a positive device time for every model, longer
    for 16x the particles; sizes outside its range are rejected."""
    params, ctrl, meas, _, _, dt, _, _ = _model_case(model, 64, 3)
    fr = ea.build_frame(model, params, ctrl, meas, dt)
    small = ea.time_hook(engine, model, fr, 1 << 16, 5)
    big = ea.time_hook(engine, model, fr, 1 << 20, 5)
    assert np.isfinite(small) and small > 0.0
    assert big > small
    us = C.c_float()
    for n, reps in [(3, 5), (0, 5), (1 << 16, 0), ((1 << 32) + 2, 5)]:
        assert ea.lib().esth_engine_time_hook(engine.h, model, ea.ptr(fr), n, reps, C.byref(us)) == ea.ESTH_ERR_INVALID
    assert ea.lib().esth_engine_time_hook(engine.h, 99, ea.ptr(fr), 1 << 16, 5, C.byref(us)) == ea.ESTH_ERR_INVALID
