"""CPU tests that pin the oracle itself: against the reference's own MTGP32
library output (golden fixtures), against closed forms, brute force and
invariants.  No GPU."""
import ctypes as C
import itertools
import os

import numpy as np
import pytest

import oracle as O

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

ESTH_TOPO_RING, ESTH_TOPO_TORUS, ESTH_TOPO_ALL2ALL = 1, 2, 3


# ---------------------------------------------------------------- MTGP32
def test_param_table_shape_and_mask(mtgp_table):
    assert mtgp_table.shape == (200, 52)
    assert set(mtgp_table[:, 3].tolist()) == {0xFFF80000}   # one mask for all sets (F4)
    # 256 words per phase need x[n-351+pos] from an earlier phase: pos <= 351 - 256
    assert np.all((mtgp_table[:, 0] >= 1) & (mtgp_table[:, 0] <= 95))


def test_seeding_matches_reference(golden, mtgp_table):
    for s in golden["seeds"]:
        s = int(s)
        for k, i in enumerate(golden["sets_init"]):
            st = O.mtgp_seed(mtgp_table, int(i), (s + int(i) + 1) & 0xFFFFFFFF)
            assert np.array_equal(st, golden[f"init_{s}"][k])


@pytest.mark.parametrize("seed", [1234, 42, 7, 99, 0xF000000D])
def test_float_words_all_sets(golden, mtgp_table, seed):
    for i in range(200):
        st = O.mtgp_seed(mtgp_table, i, (seed + i + 1) & 0xFFFFFFFF)
        assert np.array_equal(O.mtgp_generate(mtgp_table, i, st, 16, True), golden[f"f32_{seed}"][i])


def test_uint32_words_and_continuation(golden, mtgp_table):
    for i in range(4):
        st = O.mtgp_seed(mtgp_table, i, 1234 + i + 1)
        assert np.array_equal(O.mtgp_generate(mtgp_table, i, st, 4096, False), golden["u32_1234"][i])
        st = O.mtgp_seed(mtgp_table, i, 1234 + i + 1)
        parts = [O.mtgp_generate(mtgp_table, i, st, n, True) for n in (1, 767, 768, 2560)]
        assert np.array_equal(np.concatenate(parts), golden["f32long_1234"][i])


def test_reference_buffer_layout_16x256(golden, mtgp_table):
    status = np.concatenate([O.mtgp_seed(mtgp_table, i, 1234 + i + 1) for i in range(96)])
    buf = O.mtgp_fill(mtgp_table, 96, status, 768)
    assert np.array_equal(buf.reshape(96, 768), golden["buf16x256_step1"])


def test_float_words_in_unit_interval(golden):
    f = golden["f32long_1234"].view(np.float32)
    assert np.all((f >= 1.0) & (f < 2.0))


# ------------------------------------------------------------------ math
def ulps(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    sp = np.spacing(np.abs(b).astype(np.float32)).astype(np.float64)
    return np.abs(a - b) / sp


def test_expf_error_bound():
    L = O.lib()
    x = np.concatenate([np.linspace(-87.3, 88.7, 20001), np.linspace(-1, 1, 2001)]).astype(np.float32)
    got = np.array([L.or_expf(v) for v in x], np.float32)
    ref = np.exp(x.astype(np.float64))
    assert np.max(ulps(got, ref)) <= 2.0
    assert L.or_expf(-100.0) == 0.0


def test_logf_error_bound():
    L = O.lib()
    x = np.exp(np.linspace(-87, 88, 20001)).astype(np.float32)
    x = np.concatenate([x, np.linspace(0.5, 2.0, 4001, dtype=np.float32), [np.float32(2 ** -23)]])
    got = np.array([L.or_logf(v) for v in x], np.float32)
    ref = np.log(x.astype(np.float64))
    err = np.abs(got - ref)
    assert np.all(err <= np.maximum(3 * np.spacing(np.abs(ref).astype(np.float32)), 2e-7))


def test_sincosf_error_bound():
    L = O.lib()
    x = np.concatenate([np.linspace(-7, 7, 20001), np.linspace(-1000, 1000, 20001)]).astype(np.float32)
    s, c = C.c_float(), C.c_float()
    gs, gc = [], []
    for v in x:
        L.or_sincosf(v, C.byref(s), C.byref(c))
        gs.append(s.value)
        gc.append(c.value)
    xd = x.astype(np.float64)
    assert np.max(np.abs(np.array(gs) - np.sin(xd))) < 1.5e-7
    assert np.max(np.abs(np.array(gc) - np.cos(xd))) < 1.5e-7


def test_atan2f_error_bound():
    L = O.lib()
    rng = np.random.default_rng(0)
    y = rng.normal(size=20000).astype(np.float32)
    x = rng.normal(size=20000).astype(np.float32)
    got = np.array([L.or_atan2f(a, b) for a, b in zip(y, x)])
    ref = np.arctan2(y.astype(np.float64), x.astype(np.float64))
    assert np.max(np.abs(got - ref)) < 3e-7
    for a, b in [(0.0, 1.0), (1.0, 0.0), (0.0, -1.0), (-1.0, 0.0), (1.0, 1.0), (-1.0, -1.0)]:
        assert abs(L.or_atan2f(a, b) - np.arctan2(a, b)) < 3e-7


def test_rcpf_within_one_ulp():
    L = O.lib()
    rng = np.random.default_rng(2)
    x = np.concatenate([np.exp(rng.uniform(-60, 60, 20000)), [1.0, 2.0, 3.0, 1e-30, 0.999999]]).astype(np.float32)
    got = np.array([L.or_rcpf(v) for v in x], np.float32)
    ref = (1.0 / x.astype(np.float64)).astype(np.float32)
    assert np.max(np.abs(got.view(np.int32).astype(np.int64) - ref.view(np.int32).astype(np.int64))) <= 1
    assert L.or_rcpf(1.0) == 1.0 and L.or_rcpf(2.0) == 0.5


def test_polar_abs_error_bound():
    import ctypes as C
    L = O.lib()
    rng = np.random.default_rng(3)
    u = (rng.normal(size=20000) * np.exp(rng.uniform(-3, 5, 20000))).astype(np.float32)
    v = (rng.normal(size=20000) * np.exp(rng.uniform(-3, 5, 20000))).astype(np.float32)
    pts = list(zip(u, v)) + [(0.0, 0.0), (1.0, 0.0), (-1.0, 0.0), (0.0, 2.0), (0.0, -2.0), (3.0, 3.0),
                             (-3.0, -3.0), (1e-20, -1e-20)]
    r, a = C.c_float(), C.c_float()
    for uu, vv in pts:
        L.or_polar_abs(float(uu), float(vv), C.byref(r), C.byref(a))
        h = np.hypot(np.float64(uu), np.float64(vv))
        assert abs(r.value - h) <= 3.6e-7 * h + 1e-37, (uu, vv, r.value, h)
        assert abs(a.value - abs(np.arctan2(np.float64(vv), np.float64(uu)))) < 3e-7, (uu, vv)
    L.or_polar_abs(0.0, 0.0, C.byref(r), C.byref(a))
    assert r.value == 0.0 and a.value == 0.0


def test_unicycle10_model_against_fp64_twin():
    """The FP32 hook (particle_filter.cu:313-359 restated) against the
    double-precision twin written from the same definition with libm, on the
    reference's own input rows; and the straight-line limit at w = 0."""
    import os
    import esthera_amd as ea
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "unicycle_input.txt")
    meas, ctrl, dts, truth = ea.load_unicycle10_trace(path)
    p = ea.unicycle10_params_default()
    rng = np.random.default_rng(4)
    N = 2048
    for k in (1, 40, 100):
        st = np.concatenate([truth[k - 1, c] + rng.normal(0, 0.3, N) for c in range(3)]).astype(np.float32)
        normals = rng.normal(size=3 * N).astype(np.float32)
        fr = O.build_frame(3, p, ctrl[k], meas[k], float(dts[k]))
        s32, w32 = O.sample_weight(3, st, N, normals, fr, 1)
        s64, lw64 = O.sample_weight_f64(3, st, N, normals, p, ctrl[k], meas[k], float(dts[k]), 3)
        np.testing.assert_allclose(s32, s64, rtol=2e-6, atol=4e-6)
        np.testing.assert_allclose(w32, lw64, rtol=1e-4, atol=1e-3)
    # w = 0 exactly: the particle moves straight along theta
    c0 = ea.UnicycleControl(2.0, 0.0)
    st = np.array([1.0, 2.0, 0.5], np.float32)
    fr = O.build_frame(3, p, c0, meas[0], 0.5)
    s32, _ = O.sample_weight(3, st, 1, np.zeros(3, np.float32), fr, 1)
    np.testing.assert_allclose(s32, [1.0 + np.cos(0.5), 2.0 + np.sin(0.5), 0.5], rtol=1e-6)


def test_camera_model_against_fp64_twin():
    """The FP32 hook of robotarm5_camera against the double twin (projection
    written in the Brown-Conrady form with libm) on a seeded camera trace."""
    import esthera_amd as ea
    p = ea.camera_params_default()
    ctrl, meas, truth = ea.robotarm5_camera_scenario(11, 60, 0.1)
    rng = np.random.default_rng(6)
    N = 2048
    M = ea.MODEL_ROBOTARM5_CAMERA
    for k in (1, 30, 59):
        st = np.concatenate([truth[k - 1, c] + rng.normal(0, 0.05, N) for c in range(9)]).astype(np.float32)
        normals = rng.normal(size=9 * N).astype(np.float32)
        fr = O.build_frame(M, p, ctrl[k - 1], meas[k - 1], 0.1)
        s32, w32 = O.sample_weight(M, st, N, normals, fr, 1)
        s64, lw64 = O.sample_weight_f64(M, st, N, normals, p, ctrl[k - 1], meas[k - 1], 0.1, 9)
        np.testing.assert_allclose(s32, s64, rtol=2e-6, atol=4e-6)
        np.testing.assert_allclose(w32, lw64, rtol=2e-4, atol=2e-3)


def _camera_lw(p, angles, x, y, meas_xy, meas_angles):
    """log-weight of one noise-free particle at dt = 0 (the state is kept)."""
    import esthera_amd as ea
    M = ea.MODEL_ROBOTARM5_CAMERA
    c = ea.RobotArm5Control()
    m = ea.RobotArm5Measurement()
    for i in range(5):
        m.angles[i] = meas_angles[i]
    m.x, m.y = meas_xy
    fr = O.build_frame(M, p, c, m, 0.0)
    st = np.array(list(angles) + [x, y, 0.0, 0.0], np.float32)
    s32, w32 = O.sample_weight(M, st, 1, np.zeros(9, np.float32), fr, 1)
    np.testing.assert_array_equal(s32, st)
    return float(w32[0])


def test_camera_optical_axis_maps_to_principal_point():
    """All joints at 0 and the base at the origin put the point on the optical
    axis (X = Y = 0, Z = -12): its pixel is the principal point whatever the
    distortion, so a measurement there costs nothing and one (3, 4) px away
    costs 0.1 (9 + 16)."""
    import esthera_amd as ea
    p = ea.camera_params_default()
    cc = (p.cc[0], p.cc[1])
    assert _camera_lw(p, [0.0] * 5, 0.0, 0.0, cc, [0.0] * 5) == 0.0
    lw = _camera_lw(p, [0.0] * 5, 0.0, 0.0, (cc[0] + 3.0, cc[1] + 4.0), [0.0] * 5)
    np.testing.assert_allclose(lw, -0.1 * 25.0, rtol=1e-5)
    # angle terms: 0.1 per squared radian, independent of the camera
    lw = _camera_lw(p, [0.0] * 5, 0.0, 0.0, cc, [0.5, 0.0, 0.0, 0.0, -1.0])
    np.testing.assert_allclose(lw, -0.1 * 1.25, rtol=1e-5)


@pytest.mark.parametrize("xy", [(0.6, -0.4), (-0.9, 0.8), (1.5, 2.0)])
def test_camera_projection_matches_bouguet_form(xy):
    """With all joints at 0 the camera point is (x, y, -12).  The expected
    pixel is the reference's literal expression (tangential x term
    kc4 (3 xn^2 + yn^2), pf_model_soa_packed_code.cuh:119-131) in double: a
    measurement 2 px off it in u and 1 px in v costs 0.1 (4 + 1); a swap of
    kc3 / kc4 or fc_x / fc_y would move the prediction by pixels."""
    import esthera_amd as ea
    p = ea.camera_params_default()
    x, y = xy
    xn, yn = x / -12.0, y / -12.0
    k1, k2, k3, k4 = (float(v) for v in p.kc)
    r2 = xn * xn + yn * yn
    dxx = 2 * k3 * xn * yn + k4 * (3 * xn * xn + yn * yn)
    dxy = k3 * (3 * yn * yn + xn * xn) + 2 * k4 * xn * yn
    u = p.fc[0] * ((1 + k1 * r2 + k2 * r2 * r2) * xn + dxx) + p.cc[0]
    v = p.fc[1] * ((1 + k1 * r2 + k2 * r2 * r2) * yn + dxy) + p.cc[1]
    lw = _camera_lw(p, [0.0] * 5, x, y, (u + 2.0, v - 1.0), [0.0] * 5)
    np.testing.assert_allclose(lw, -0.1 * 5.0, rtol=2e-3)
    # the pinhole limit: no distortion, unit focal length, no offset
    q = ea.CameraParams()
    q.fc[0] = q.fc[1] = 1.0
    lw = _camera_lw(q, [0.0] * 5, x, y, (xn, yn + 0.5), [0.0] * 5)
    np.testing.assert_allclose(lw, -0.1 * 0.25, rtol=1e-4)


def test_camera_nan_cost_is_zero_weight():
    """Base at the origin, joints (0, 0, pi, 0, pi): joint 2 turns the first
    link's offset back onto the base's plane and joint 4 the third's, so the
    point reaches Z = 0 with X = 0, and X / Z is 0 / 0.  Its weight is 0
    (log -inf), never NaN."""
    import esthera_amd as ea
    p = ea.camera_params_default()
    a = [0.0, 0.0, np.pi, 0.0, np.pi]
    lw = _camera_lw(p, a, 0.0, 0.0, (p.cc[0], p.cc[1]), a)
    assert lw == -np.inf


def _carry_case(threshold, weights):
    """Zero process noise and one common initial state: every particle follows
    the same deterministic path, so every weight is the same function of the
    steps (no sort, exchange or resampling can change any particle)."""
    import ctypes as C
    import esthera_amd as ea
    p = ea.UnicycleParams()
    p.num_landmarks = 4
    p.sigma_x = p.sigma_y = p.sigma_theta = 0.0
    p.sigma_range, p.sigma_bearing = 0.5, 0.1
    for j, (x, y) in enumerate([(5.0, 1.0), (-3.0, 4.0), (2.0, -6.0), (7.0, 7.0)]):
        p.lx[j], p.ly[j] = x, y
    ctrl = ea.UnicycleControl(1.0, 0.1)
    meas = ea.UnicycleMeasurement()
    for j in range(4):
        meas.range[j], meas.bearing[j] = 5.0 + j, 0.3 * j - 0.4
    d = ea.filter_desc(ea.MODEL_UNICYCLE, m=64, G=4, t=1, topology=ea.TOPO_RING, threshold=threshold,
                       weights=weights, seed=5, carry=1)
    f = O.Filter(d, p, C.sizeof(p))
    f.init_uniform(np.zeros(3, np.float32))
    # FP64 path of the common particle
    x = np.zeros(3)
    lls = []
    for k in range(4):
        f.step(ctrl, meas, 0.1)
        x = np.array([x[0] + 0.1 * np.cos(x[2]), x[1] + 0.1 * np.sin(x[2]), x[2] + 0.01])
        acc = 0.0
        for j in range(4):
            dx, dy = p.lx[j] - x[0], p.ly[j] - x[1]
            er = (meas.range[j] - np.hypot(dx, dy)) / 0.5
            eb = np.angle(np.exp(1j * (meas.bearing[j] - (np.arctan2(dy, dx) - x[2])))) / 0.1
            acc += er * er + eb * eb
        lls.append(-0.5 * acc)
    st, w = f.read(3)
    return np.array(lls), st.reshape(3, -1), w


def test_carried_weights_accumulate_without_resampling():
    lls, st, w = _carry_case(-1.0, 1)     # the gate u - 1 > -1 skips every sub-filter
    assert np.allclose(st, st[:, :1], atol=0)          # still one common particle
    np.testing.assert_allclose(w, lls.sum(), rtol=2e-5)
    lls, st, w = _carry_case(-1.0, 0)     # linear: the product of the likelihoods
    np.testing.assert_allclose(w, np.exp(lls.sum()), rtol=1e-4)


def test_carried_weights_restart_after_resampling():
    lls, st, w = _carry_case(1.0, 1)      # every sub-filter resampled: weights restart every step
    np.testing.assert_allclose(w, 0.0, atol=0)
    lls, st, w = _carry_case(1.0, 0)
    np.testing.assert_allclose(w, 1.0, atol=0)


def test_boxmuller_closed_form():
    rng = np.random.default_rng(1)
    u = (1.0 + rng.random(10000, dtype=np.float32)).astype(np.float32)
    z = O.boxmuller(u)
    u1 = 2.0 - u[0::2].astype(np.float64)
    u2 = 2.0 - u[1::2].astype(np.float64)
    r = np.sqrt(-2 * np.log(u1))
    assert np.max(np.abs(z[0::2] - r * np.cos(2 * np.pi * u2))) < 2e-6
    assert np.max(np.abs(z[1::2] - r * np.sin(2 * np.pi * u2))) < 2e-6
    # u1 in (0, 1]: the largest input word maps to the smallest u1 without log(0)
    zz = O.boxmuller(np.array([np.float32(2 - 2 ** -23), 1.25], np.float32))
    assert np.all(np.isfinite(zz))


# ------------------------------------------------------------------ sort
@pytest.mark.parametrize("m", [2, 4, 8])
def test_bitonic_brute_force(m):
    """Every key pattern over {0,1,2}^m: sorted descending, a permutation, and
    the gathered payload follows the permutation."""
    pats = np.array(list(itertools.product(range(3), repeat=m)), np.float32)
    G = len(pats)
    w = pats.ravel()
    payload = np.arange(w.size, dtype=np.float32)
    sw, out, win, perm = O.block_sort(w, payload, 1, m, G)
    sw = sw.reshape(G, m)
    perm = perm.reshape(G, m)
    assert np.array_equal(sw, -np.sort(-pats, axis=1))
    assert np.all(np.sort(perm, axis=1) == np.arange(m))
    assert np.array_equal(np.take_along_axis(pats, perm, 1), sw)
    assert np.array_equal(out.reshape(G, m), perm + np.arange(G)[:, None] * m)
    assert np.array_equal(win[0::2], sw[:, 0]) and np.array_equal(win[1::2], np.arange(G))


def test_bitonic_ties_follow_the_network():
    """Equal keys swap under ddd == 0 (bitonic_sort.cl:16): on all-equal keys the
    network is a fixed permutation, not the identity."""
    m = 8
    _, _, _, perm = O.block_sort(np.zeros(m, np.float32), np.zeros(m, np.float32), 1, m, 1)
    assert sorted(perm.tolist()) == list(range(m))
    assert perm.tolist() != list(range(m))


def test_bitonic_large_random():
    rng = np.random.default_rng(2)
    for m in (256, 512, 1024):
        w = rng.random(m * 3, dtype=np.float32)
        sw, _, _, perm = O.block_sort(w, np.zeros(w.size, np.float32), 1, m, 3)
        for g in range(3):
            blk = w[g * m:(g + 1) * m]
            assert np.array_equal(sw[g * m:(g + 1) * m], np.sort(blk)[::-1])
            assert np.array_equal(blk[perm[g * m:(g + 1) * m]], sw[g * m:(g + 1) * m])


# ---------------------------------------------------------------- argmax
def test_argmax_distinct_and_ties():
    rng = np.random.default_rng(3)
    G = 1000
    win = np.zeros(2 * G, np.float32)
    win[0::2] = rng.permutation(G).astype(np.float32)
    win[1::2] = np.arange(G)
    r = O.argmax(win, G, 0)
    assert r[1] == np.argmax(win[0::2]) and r[0] == G - 1
    # all equal: the strided pass keeps each lane's last element, the tree
    # prefers `other` (higher lane), so the winner is the last element of lane 127
    win[0::2] = 1.0
    r = O.argmax(win, G, 0)
    last_of_lane127 = max(i for i in range(G) if i % 128 == 127)
    assert r[1] == last_of_lane127
    # linear mode starts at (-1,-1): all weights below -1 never win
    small = np.array([-2.0, 0.0, -3.0, 1.0], np.float32)
    assert O.argmax(small, 2, 0)[1] == -1.0
    assert O.argmax(small, 2, 1)[1] == 0.0


# -------------------------------------------------------------- exchange
def sender_side_exchange(state, w, S, m, G, t, topo, dimx=0, dimy=0):
    """Sender-side formulation (each block writes its best t into its
    neighbours' worst slots, as exchange_ring / exchange_2dtorus do) from a
    snapshot; the oracle implements the receiver side."""
    st = state.reshape(S, G, m).copy()
    ww = w.reshape(G, m).copy()
    src_st = st.copy()
    src_w = ww.copy()
    for b in range(G):
        for k in range(t):
            if topo == ESTH_TOPO_RING:
                dest = [((b + 1) % G, m - 1 - k)]
            else:
                x, y = b % dimx, b // dimx
                left = y * dimx + (x - 1) % dimx
                right = y * dimx + (x + 1) % dimx
                top = ((y - 1) % dimy) * dimx + x
                bottom = ((y + 1) % dimy) * dimx + x
                dest = [(left, m - 1 - k), (right, m - 1 - k - t), (top, m - 1 - k - 2 * t),
                        (bottom, m - 1 - k - 3 * t)]
            for blk, slot in dest:
                ww[blk, slot] = src_w[b, k]
                st[:, blk, slot] = src_st[:, b, k]
    return st.ravel(), ww.ravel()


@pytest.mark.parametrize("topo,m,G,t,dx,dy", [(ESTH_TOPO_RING, 16, 5, 3, 0, 0), (ESTH_TOPO_TORUS, 16, 16, 2, 4, 4),
                                              (ESTH_TOPO_TORUS, 32, 32, 4, 4, 8), (ESTH_TOPO_TORUS, 8, 6, 1, 2, 3)])
def test_exchange_matches_sender_side_map(topo, m, G, t, dx, dy):
    rng = np.random.default_rng(4)
    S = 2
    w = -np.sort(-rng.random((G, m), dtype=np.float32), axis=1).ravel()
    st = rng.normal(size=S * m * G).astype(np.float32)
    ost, ow = O.exchange(st, w, S, m, G, t, topo, dx, dy)
    rst, rw = sender_side_exchange(st, w, S, m, G, t, topo, dx, dy)
    assert np.array_equal(ow, rw)
    assert np.array_equal(ost, rst)


def test_exchange_all2all_is_global_top_t():
    rng = np.random.default_rng(5)
    m, G, t, S = 16, 10, 3, 1
    w = -np.sort(-rng.integers(0, 20, (G, m)).astype(np.float32), axis=1)
    st = np.arange(m * G, dtype=np.float32)
    ost, ow = O.exchange(st, w.ravel(), S, m, G, t, ESTH_TOPO_ALL2ALL)
    cand = [(w[g, k], g * t + k, g, k) for g in range(G) for k in range(t)]
    cand.sort(key=lambda c: (-c[0], c[1]))
    for b in range(G):
        for k in range(t):
            g, kk = cand[k][2], cand[k][3]
            assert ow[b * m + m - 1 - k] == w[g, kk]
            assert ost[b * m + m - 1 - k] == g * m + kk


# ------------------------------------------------------------ resampling
def test_scan_selection_closed_form():
    """Integer weights: the Blelloch scan is exact, so the selection must equal
    the largest i with cdf[i] <= u*total (numpy's searchsorted)."""
    rng = np.random.default_rng(6)
    m, G, S = 64, 20, 1
    w = rng.integers(0, 4, m * G).astype(np.float32)
    rnd = (1.0 + rng.random((m + 1) * G, dtype=np.float32)).astype(np.float32)
    src = np.arange(m * G, dtype=np.float32)
    _, sel = O.resample(0, np.zeros(m * G, np.float32), src, S, w, rnd, m, G, 1.0, 0, 0)
    for g in range(G):
        wb = w[g * m:(g + 1) * m].astype(np.float64)
        cdf = np.concatenate([[0.0], np.cumsum(wb)[:-1]])
        total = wb.sum()
        for i in range(m):
            s = np.float32(np.float32(rnd[g * m + i] - np.float32(1.0)) * np.float32(total))
            assert sel[g * m + i] == np.searchsorted(cdf, s, side="right") - 1


def test_systematic_offspring_counts():
    """Systematic selection: offspring counts differ from m*p_i by less than 1
    and sum to m."""
    rng = np.random.default_rng(7)
    m, G = 256, 8
    w = (rng.random(m * G, dtype=np.float32) ** 3).astype(np.float32)
    rnd = (1.0 + rng.random((m + 1) * G, dtype=np.float32)).astype(np.float32)
    _, sel = O.resample(1, np.zeros(m * G, np.float32), np.zeros(m * G, np.float32), 1, w, rnd, m, G, 1.0, 0, 0)
    for g in range(G):
        counts = np.bincount(sel[g * m:(g + 1) * m], minlength=m)
        p = w[g * m:(g + 1) * m].astype(np.float64)
        p /= p.sum()
        assert counts.sum() == m
        assert np.all(np.abs(counts - m * p) < 1.0 + 1e-3)
        assert np.all(np.diff(sel[g * m:(g + 1) * m]) >= 0)


def test_gate_skips_subfilters():
    m, G = 8, 4
    w = np.ones(m * G, np.float32)
    rnd = np.full((m + 1) * G, 1.5, np.float32)
    rnd[m * G + 1] = 1.9      # sub-filter 1's gate uniform 0.9 > 0.5: not resampled
    dst = np.full(m * G, -7.0, np.float32)
    d, sel = O.resample(0, dst, np.arange(m * G, dtype=np.float32), 1, w, rnd, m, G, 0.5, 0, 0)
    assert np.all(sel[m:2 * m] == -1) and np.all(d[m:2 * m] == -7.0)
    assert np.all(sel[:m] >= 0)


@pytest.mark.parametrize("m", [2, 16, 256, 1024])
def test_alias_invariant(m):
    rng = np.random.default_rng(m)
    G = 5
    w = (rng.random(m * G, dtype=np.float32) ** 4).astype(np.float32)
    prob, alias = O.alias_table(w, m, G, 0)
    for g in range(G):
        p = w[g * m:(g + 1) * m].astype(np.float64)
        p /= p.sum()
        pr = np.clip(prob[g * m:(g + 1) * m].astype(np.float64), 0, 1)
        mass = pr.copy()
        np.add.at(mass, alias[g * m:(g + 1) * m], 1.0 - pr)
        np.testing.assert_allclose(mass, m * p, atol=2e-5 * m)


def test_alias_draw_distribution():
    """chi-square of 256k alias draws against the normalised weights."""
    rng = np.random.default_rng(8)
    m, G = 16, 16384
    wb = np.array([1, 2, 3, 4, 5, 6, 7, 8, 0, 0, 1, 1, 2, 2, 10, 0.5], np.float32)
    w = np.tile(wb, G)
    rnd = (1.0 + rng.random((2 * m + 1) * G, dtype=np.float32)).astype(np.float32)
    _, sel = O.resample(2, np.zeros(m * G, np.float32), np.zeros(m * G, np.float32), 1, w, rnd, m, G, 1.0, 0, 0)
    counts = np.bincount(sel, minlength=m).astype(np.float64)
    p = wb / wb.sum()
    exp = p * sel.size
    nz = exp > 0
    chi2 = np.sum((counts[nz] - exp[nz]) ** 2 / exp[nz])
    assert chi2 < 40.0          # 13 dof: p ~ 2e-4
    assert np.all(counts[~nz] == 0)


def test_cpu_baseline_build_gives_the_same_results(tmp_path):
    """The CPU baseline's -O3 -march=native build of the oracle (make -C oracle
    baseline) steps a filter bit-identically to the checker's build: the flags
    change speed, not arithmetic (no contraction, no reassociation)."""
    import ctypes
    import subprocess
    import esthera_amd as ea
    r = subprocess.run(["make", "-s", "-C", os.path.join(REPO, "oracle"), "baseline", f"OUT={tmp_path}"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    desc = ea.filter_desc(ea.MODEL_UNICYCLE, m=256, G=16, t=1, topology=ea.TOPO_TORUS,
                          resampler=ea.RESAMPLE_SYSTEMATIC, weights=ea.WEIGHTS_LOG, seed=1234)
    p, ctrl, meas, truth = ea.unicycle_scenario(1234, 32, 50.0, 3)
    out = []
    for path in (O.LIB_PATH, str(tmp_path / "liboracle_native.so")):
        O.use_library(path)
        f = O.Filter(desc, p, ctypes.sizeof(p))
        f.init_uniform(np.zeros(3, np.float32))
        for k in range(3):
            f.step(ctrl[k], meas[k], 0.1)
        out.append(f.read(3))
        del f
    O.use_library(O.LIB_PATH)
    assert np.array_equal(out[0][0].view(np.uint32), out[1][0].view(np.uint32))
    assert np.array_equal(out[0][1].view(np.uint32), out[1][1].view(np.uint32))
