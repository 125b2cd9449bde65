"""The model list: every listed model is reachable by name and id, and the
randomwalk3 model (the example of adding one) is pinned against closed forms.

randomwalk3 is linear-Gaussian, so its filtering posterior has a closed form
(the Kalman filter): the oracle's whole filter on it must track that mean to
Monte Carlo accuracy, which checks the oracle's hook, weights, resampling and
posterior together against mathematics rather than against itself.
"""
import ctypes as C

import numpy as np
import pytest

import esthera_amd as ea
import oracle as O

NAMES = {"robotarm5": ea.MODEL_ROBOTARM5, "unicycle": ea.MODEL_UNICYCLE, "dh6": ea.MODEL_DH6,
         "unicycle10": ea.MODEL_UNICYCLE10, "robotarm5_camera": ea.MODEL_ROBOTARM5_CAMERA,
         "randomwalk3": ea.MODEL_RANDOMWALK3}


def test_every_model_listed():
    L = ea.lib()
    S, V, ps, cs, ms = C.c_int(), C.c_int(), C.c_size_t(), C.c_size_t(), C.c_size_t()
    for name, model in NAMES.items():
        assert L.esth_model_lookup(name.encode()) == model
        ea.check(L.esth_model_info(model, C.byref(S), C.byref(V), C.byref(ps), C.byref(cs), C.byref(ms)))
        assert (S.value, V.value) == (ea.STATE_DIM[model], ea.NOISE_DIM[model])
        pt = ea.PARAMS_TYPE[model]
        assert ps.value == (0 if pt is None else C.sizeof(pt))
        assert cs.value == C.sizeof(ea.CONTROL_TYPE[model]) and ms.value == C.sizeof(ea.MEAS_TYPE[model])
    assert L.esth_model_info(len(NAMES), None, None, None, None, None) == ea.ESTH_ERR_INVALID


def rw3_inputs(seed=3):
    p, ctrl, meas, truth = ea.randomwalk3_scenario(seed, 6)
    return p, ctrl, meas, truth


def test_randomwalk3_frame():
    p, ctrl, meas, _ = rw3_inputs()
    fr = ea.build_frame(ea.MODEL_RANDOMWALK3, p, ctrl[0], meas[0], 0.1)
    assert np.array_equal(fr.view(np.uint32), O.build_frame(ea.MODEL_RANDOMWALK3, p, ctrl[0], meas[0], 0.1)
                          .view(np.uint32))
    v = np.array(ctrl[0].v, np.float32)
    assert np.array_equal(fr[1:4], v * np.float32(0.1))
    assert fr[4] == np.float32(p.sigma_q) and fr[5] == np.float32(0.5) / np.float32(p.sigma_r * p.sigma_r)
    assert np.array_equal(fr[6:9], np.array(meas[0].p, np.float32))
    with pytest.raises(ea.EsthError):   # sigma_r must be positive
        ea.build_frame(ea.MODEL_RANDOMWALK3, ea.RandomWalk3Params(0.1, 0.0), ctrl[0], meas[0], 0.1)


def test_randomwalk3_oracle_hook_closed_form():
    """FP32 hook and FP64 twin against p' = p + v dt + sigma_q z and
    log w = -|m - p'|^2 / (2 sigma_r^2), written here in float64."""
    p, ctrl, meas, _ = rw3_inputs()
    N, dt = 1000, 0.1
    rng = np.random.default_rng(5)
    st = rng.normal(0.0, 1.0, 3 * N).astype(np.float32)
    z = rng.normal(0.0, 1.0, 3 * N).astype(np.float32)
    fr = O.build_frame(ea.MODEL_RANDOMWALK3, p, ctrl[0], meas[0], dt)
    s32, lw32 = O.sample_weight(ea.MODEL_RANDOMWALK3, st, N, z, fr, ea.WEIGHTS_LOG)
    s64, lw64 = O.sample_weight_f64(ea.MODEL_RANDOMWALK3, st, N, z, p, ctrl[0], meas[0], dt, 3)
    v = np.array(ctrl[0].v, np.float64)[:, None]
    m = np.array(meas[0].p, np.float64)[:, None]
    x = st.astype(np.float64).reshape(3, N) + v * np.float64(np.float32(dt)) + \
        np.float64(np.float32(p.sigma_q)) * z.astype(np.float64).reshape(3, N)
    lw = -0.5 * np.sum((m - x) ** 2, axis=0) / np.float64(np.float32(p.sigma_r)) ** 2
    np.testing.assert_allclose(s64.reshape(3, N), x, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(lw64, lw, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(s32.reshape(3, N), x, rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(lw32, lw, rtol=1e-5, atol=1e-4)


def kalman_means(p, ctrl, meas, dt, steps):
    """Posterior means of p_k = p_{k-1} + v dt + N(0, sigma_q^2), m_k = p_k +
    N(0, sigma_r^2) from a point prior at the origin (per axis, independent)."""
    q, r = np.float64(np.float32(p.sigma_q)) ** 2, np.float64(np.float32(p.sigma_r)) ** 2
    x, P, out = np.zeros(3), 0.0, []
    for k in range(steps):
        x = x + np.array(ctrl[k].v, np.float64) * np.float64(np.float32(dt))
        P = P + q
        K = P / (P + r)
        x = x + K * (np.array(meas[k].p, np.float64) - x)
        P = (1.0 - K) * P
        out.append((x.copy(), P))
    return out


@pytest.mark.parametrize("resampler", [ea.RESAMPLE_SYSTEMATIC, ea.RESAMPLE_ALIAS])
def test_randomwalk3_oracle_filter_tracks_kalman(resampler):
    """The oracle's whole filter (65,536 particles, ring exchange) against
    the Kalman posterior: per step the mean lies within 24 standard errors
    sqrt(P / N) (resampled particles are correlated, so the effective N is
    below N); a dropped v dt or sigma term misses by 1e-2 or more."""
    p, ctrl, meas, _ = rw3_inputs(11)
    steps, dt = 6, 0.1
    desc = ea.filter_desc(ea.MODEL_RANDOMWALK3, m=512, G=128, t=1, topology=ea.TOPO_RING, resampler=resampler,
                          weights=ea.WEIGHTS_LOG, seed=9)
    o = O.Filter(desc, p, C.sizeof(p))
    o.init_uniform(np.zeros(3, np.float32))
    N = 512 * 128
    for k, (x, P) in enumerate(kalman_means(p, ctrl, meas, dt, steps)):
        o.step(ctrl[k], meas[k], dt)
        m32, m64, _ = o.estimate(3)
        tol = 24.0 * np.sqrt(P / N)
        assert np.all(np.abs(m64 - x) < tol), (k, m64, x, tol)
        assert np.all(np.abs(m32 - m64) < 1e-4)
