"""The C ABI without a GPU: the library loads and exports every function the
header declares; host-side entry points (models, frames, MTGP seeding and
parameter table, scenarios, trace reading, autotuner choice) behave."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import esthera_amd as ea
import oracle as O

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(REPO, "include", "esthera.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*[\w\s\*]+?\b(esth_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_function():
    lib = C.CDLL(ea.LIB_PATH)
    names = header_functions()
    assert len(names) > 60
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_models_and_defaults():
    L = ea.lib()
    assert L.esth_model_lookup(b"robotarm5") == 0 and L.esth_model_lookup(b"unicycle") == 1
    assert L.esth_model_lookup(b"dh6") == 2 and L.esth_model_lookup(b"nope") == -1
    S, V, ps, cs, ms = C.c_int(), C.c_int(), C.c_size_t(), C.c_size_t(), C.c_size_t()
    assert L.esth_model_lookup(b"unicycle10") == 3
    for model, (s, v) in {0: (9, 9), 1: (3, 3), 2: (12, 6), 3: (3, 3)}.items():
        ea.check(L.esth_model_info(model, C.byref(S), C.byref(V), C.byref(ps), C.byref(cs), C.byref(ms)))
        assert (S.value, V.value) == (s, v)
        assert ms.value == C.sizeof(ea.MEAS_TYPE[model])
    d = ea.FilterDesc()
    L.esth_filter_desc_default(C.byref(d), 0)
    assert (d.num_particles, d.num_blocks, d.num_transfer) == (256, 16, 1)   # pf_generic.c:1072-1079
    assert d.resampling_threshold == 1.0 and d.topology == ea.TOPO_RING


def test_camera_model_registered():
    L = ea.lib()
    assert L.esth_model_lookup(b"robotarm5_camera") == ea.MODEL_ROBOTARM5_CAMERA
    S, V, ps, cs, ms = C.c_int(), C.c_int(), C.c_size_t(), C.c_size_t(), C.c_size_t()
    ea.check(L.esth_model_info(ea.MODEL_ROBOTARM5_CAMERA, C.byref(S), C.byref(V), C.byref(ps), C.byref(cs),
                               C.byref(ms)))
    assert (S.value, V.value, ps.value) == (9, 9, C.sizeof(ea.CameraParams))
    assert (cs.value, ms.value) == (C.sizeof(ea.RobotArm5Control), C.sizeof(ea.RobotArm5Measurement))
    p = ea.camera_params_default()   # pf_model_soa_packed_code.cuh:33-43
    np.testing.assert_allclose(list(p.fc) + list(p.cc), [408.65831, 409.69059, 166.32611, 108.93539], rtol=1e-7)
    np.testing.assert_allclose(list(p.kc), [-0.21561, 1.60704, -0.01291, 0.00007], rtol=1e-6)
    ctrl, meas, truth = ea.robotarm5_camera_scenario(5, 3, 0.1)
    fr = ea.build_frame(ea.MODEL_ROBOTARM5_CAMERA, p, ctrl[0], meas[0], 0.1)
    assert np.array_equal(fr.view(np.uint32),
                          O.build_frame(ea.MODEL_ROBOTARM5_CAMERA, p, ctrl[0], meas[0], 0.1).view(np.uint32))
    assert fr[11] == np.float32(meas[0].x) and fr[13] == np.float32(p.fc[0]) and fr[20] == np.float32(p.kc[3])
    with pytest.raises(ea.EsthError):   # the camera model needs its intrinsics
        ea.build_frame(ea.MODEL_ROBOTARM5_CAMERA, None, ctrl[0], meas[0], 0.1)


@pytest.mark.parametrize("S", [1, 2, 3, 5, 9, 12])
def test_state_layouts(S):
    """SOA planes, AOS structs and PACKED groups of four as the header defines
    them, built here with numpy, and the converter between every pair."""
    n = 37
    soa = np.arange(S * n, dtype=np.float32).reshape(S, n) + 0.5
    aos = soa.T.copy().reshape(-1)
    groups = []
    for q in range(0, S, 4):
        groups.append(soa[q:q + 4].T.reshape(-1))   # group q: particle-major, its w planes interleaved
    packed = np.concatenate(groups)
    by = {ea.LAYOUT_SOA: soa.reshape(-1), ea.LAYOUT_AOS: aos, ea.LAYOUT_PACKED: packed}
    for a, src in by.items():
        for b, want in by.items():
            got = ea.state_convert(src, S, n, a, b)
            assert np.array_equal(got, want), (a, b)
    if S == 9:   # groups of planes 0-3, 4-7, 8 (the reference's arm packs 0-3, 4, 5-8)
        assert packed[4 * n] == soa[4, 0] and packed[4 * n + 3] == soa[7, 0] and packed[8 * n] == soa[8, 0]
    with pytest.raises(ea.EsthError):
        ea.state_convert(soa, S, n, ea.LAYOUT_SOA, 7)


def test_frames_match_oracle():
    p, ctrl, meas, truth = ea.unicycle_scenario(42, 32, 50.0, 2)
    assert np.array_equal(ea.build_frame(1, p, ctrl[1], meas[1], 0.1).view(np.uint32),
                          O.build_frame(1, p, ctrl[1], meas[1], 0.1).view(np.uint32))
    dp, dmeas, dtruth = ea.dh6_scenario(7, 2)
    assert np.array_equal(ea.build_frame(2, dp, ea.DH6Control(), dmeas[0], 0.01).view(np.uint32),
                          O.build_frame(2, dp, ea.DH6Control(), dmeas[0], 0.01).view(np.uint32))
    with pytest.raises(ea.EsthError):
        bad = ea.UnicycleParams()
        bad.num_landmarks = 70
        ea.build_frame(1, bad, ctrl[0], meas[0], 0.1)


def test_unicycle10_trace_params_and_frame():
    path = os.path.join(REPO, "tests", "golden", "unicycle_input.txt")
    meas, ctrl, dt, truth = ea.load_unicycle10_trace(path)
    rows = [ln.split() for ln in open(path) if len(ln.split()) == 16]
    assert len(meas) == len(rows) == 131
    first = [float(x) for x in rows[0]]
    assert np.allclose(list(meas[0].range), first[:10])
    assert np.isclose(ctrl[0].v, first[10]) and np.isclose(ctrl[0].omega, first[11]) and np.isclose(dt[0], first[12])
    assert np.allclose(truth[0], first[13:16])
    p = ea.unicycle10_params_default()
    assert list(p.sx) == [12, 12, 13, 14, 16, 16, 18, 20, 20, 20]       # particle_filter.cu:69-74
    assert list(p.sy) == [12, 37, 25, 18, 12, 37, 30, 12, 25, 37]
    assert list(p.sz) == [3.0] * 10
    assert np.isclose(p.noise_v, 0.5) and np.isclose(p.noise_omega, 0.1) and np.isclose(p.noise_gamma, 0.01)
    fr = ea.build_frame(3, p, ctrl[5], meas[5], float(dt[5]))
    assert np.array_equal(fr.view(np.uint32), O.build_frame(3, p, ctrl[5], meas[5], float(dt[5])).view(np.uint32))
    assert np.isclose(fr[6], np.log(1e5) - 5 * np.log(2 * np.pi))


def test_mtgp_table_and_seeding(golden, mtgp_table):
    assert np.array_equal(ea.mtgp_params(), mtgp_table)
    for k, i in enumerate(golden["sets_init"]):
        assert np.array_equal(ea.mtgp_seed_state(int(i), 1234 + int(i) + 1), golden["init_1234"][k])


def test_unicycle_scenario_statistics():
    p, ctrl, meas, truth = ea.unicycle_scenario(42, 32, 50.0, 500)
    assert p.num_landmarks == 32
    lx, ly = np.array(p.lx[:32]), np.array(p.ly[:32])
    assert np.all(np.abs(lx) <= 50) and np.all(np.abs(ly) <= 50)
    om = np.array([c.omega for c in ctrl])
    assert np.allclose(om, 0.2 * np.sin(0.05 * np.arange(500)), atol=1e-6)
    # truth follows x += v dt cos(theta) + noise: residual std ~ 0.05
    dx = truth[1:, 0] - truth[:-1, 0] - 0.1 * np.cos(truth[:-1, 2])
    assert 0.04 < np.std(dx) < 0.06
    # range residuals ~ 0.1, bearing residuals ~ 0.02
    r = np.array([m.range[:32] for m in meas])
    rt = np.hypot(lx[None, :] - truth[:, 0:1], ly[None, :] - truth[:, 1:2])
    assert 0.09 < np.std(r - rt) < 0.11
    b = np.array([m.bearing[:32] for m in meas])
    bt = np.arctan2(ly[None, :] - truth[:, 1:2], lx[None, :] - truth[:, 0:1]) - truth[:, 2:3]
    eb = np.angle(np.exp(1j * (b - bt)))
    assert 0.018 < np.std(eb) < 0.022
    p4, _, _, _ = ea.unicycle_scenario(1234, 4, 0.0, 1)
    assert sorted(zip(p4.lx[:4], p4.ly[:4])) == sorted([(10, 10), (-10, 10), (-10, -10), (10, -10)])


def test_dh6_scenario():
    p, meas, truth = ea.dh6_scenario(7, 200)
    a, d = np.array(p.a), np.array(p.d)
    assert np.all((a >= 0.2) & (a <= 0.6)) and np.all((d >= -0.1) & (d <= 0.1))
    enc = np.array([m.q[:6] for m in meas]) - truth[1:, :6]
    assert 0.004 < np.std(enc) < 0.006


def test_trace_loader_reads_reference_trace():
    path = os.path.join(REPO, "tests", "golden", "robotarm_input5.txt")
    meas, ctrl, dt, truth = ea.load_robotarm_trace(path)
    assert len(meas) == 100
    first = [float(x) for x in open(path).readline().split()]
    assert np.isclose(meas[0].x, first[0]) and np.isclose(meas[0].angles[4], first[6])
    assert np.isclose(ctrl[0].angles[0], first[7]) and np.isclose(dt[0], first[12])
    assert np.isclose(truth[0, 5], first[13]) and np.isclose(truth[0, 0], first[15])


def test_trace_loader_stops_at_garbled_row(tmp_path):
    p = tmp_path / "t.txt"
    row = " ".join(["0.5"] * 20)
    p.write_text(row + "\n" + row + "\n" + "1 2 3 x\n")
    meas, ctrl, dt, truth = ea.load_robotarm_trace(str(p))
    assert len(meas) == 2
    with pytest.raises(ea.EsthError):
        ea.load_robotarm_trace(str(tmp_path / "missing.txt"))


def test_estimate_error_matches_reference_formula():
    est = np.arange(9, dtype=np.float32)
    act = est + np.float32(0.5)
    e = ea.lib().esth_robotarm5_estimate_error(ea.ptr(est), ea.ptr(act))
    assert np.isclose(e, 7 * 0.25 * 0.1, rtol=1e-6)   # 5 angles + x + y, each (0.5)^2 * 0.1


def test_autotune_pick_rules():
    us = np.array([5.0, 4.0, 3.0, 3.5, 3.0], np.float32)   # m = 64 .. 1024
    assert ea.autotune_pick(us, 65536, 256) == 1024          # tie 256 vs 1024: larger
    assert ea.autotune_pick(us, 64 * 3, 256) == 64           # only 64 divides 192
    assert ea.autotune_pick(us, 512 * 3, 256) == 256         # 256 cheaper than 512
    assert ea.autotune_pick(us, 128 * 5, 64) == 128
    with pytest.raises(ea.EsthError):
        ea.autotune_pick(us, 100, 64)


def test_filter_controls_reject_null_handles():
    """Filter switches fail with a status (no device call) on a null filter."""
    import ctypes as C
    L = ea.lib()
    active = C.c_int(7)
    assert L.esth_filter_set_fusion(None, 1, C.byref(active)) != 0
    assert active.value == 7
    assert L.esth_filter_set_prefetch(None, 1) != 0
    assert "null filter" in L.esth_last_error().decode()


def test_batch_autotune_decl_parses():
    assert "esth_batch_create" in header_functions() and "esth_autotune" in header_functions()


def _struct_bits(rows):
    return b"".join(bytes(r) for r in rows)


def test_trace_writers_roundtrip(tmp_path):
    """Traces written in the reference formats read back bit-identically: the
    reference's own robot-arm and unicycle inputs, and a seeded camera trace."""
    arm = os.path.join(REPO, "tests", "golden", "robotarm_input5.txt")
    meas, ctrl, dt, truth = ea.load_robotarm_trace(arm)
    out = str(tmp_path / "arm.txt")
    ea.write_robotarm_trace(out, meas, ctrl, dt, truth)
    m2, c2, d2, t2 = ea.load_robotarm_trace(out)
    assert _struct_bits(m2) == _struct_bits(meas) and _struct_bits(c2) == _struct_bits(ctrl)
    assert np.array_equal(d2.view(np.uint32), dt.view(np.uint32))
    assert np.array_equal(t2.view(np.uint32), truth.view(np.uint32))
    # the reference's file itself: same numbers, row for row
    ref = np.loadtxt(arm, dtype=np.float64).astype(np.float32)
    mine = np.loadtxt(out, dtype=np.float64).astype(np.float32)
    assert np.array_equal(mine.view(np.uint32), ref.view(np.uint32))
    uni = os.path.join(REPO, "tests", "golden", "unicycle_input.txt")
    um, uc, ud, ut = ea.load_unicycle10_trace(uni)
    out = str(tmp_path / "uni.txt")
    ea.write_unicycle10_trace(out, um, uc, ud, ut)
    um2, uc2, ud2, ut2 = ea.load_unicycle10_trace(out)
    assert _struct_bits(um2) == _struct_bits(um) and _struct_bits(uc2) == _struct_bits(uc)
    assert np.array_equal(ut2.view(np.uint32), ut.view(np.uint32))
    cc, cm, ct = ea.robotarm5_camera_scenario(8, 25, 0.1)
    out = str(tmp_path / "cam.txt")
    tr = ct[1:].copy()
    ea.write_robotarm_trace(out, cm, cc, np.full(25, 0.1, np.float32), tr)
    m3, c3, d3, t3 = ea.load_robotarm_trace(out)
    assert _struct_bits(m3) == _struct_bits(cm) and len(m3) == 25
    assert np.array_equal(t3[:, :7], tr[:, :7]) and not t3[:, 7:].any()


def _cli(*args):
    import subprocess
    return subprocess.run([ea.CLI_PATH, *args], capture_output=True, text=True, timeout=60)


def test_cli_arguments_without_gpu(tmp_path):
    """The command line's argument handling (pf_generic.c:1103-1138) up to the
    point where it would open the device."""
    r = _cli("-h")
    assert r.returncode != 0 and "Usage:" in r.stderr and "input_file" in r.stderr
    r = _cli("-m", "64")
    assert r.returncode != 0 and "no input file given" in r.stderr
    r = _cli("-c", "x")
    assert r.returncode != 0 and "no CPU device" in r.stderr
    r = _cli("-d", "1:0", "x")
    assert r.returncode != 0 and "-d" in r.stderr
    r = _cli("-M", "dh6", "x")
    assert r.returncode != 0 and "trace-driven" in r.stderr
    r = _cli("-m", "0", "x")
    assert r.returncode != 0 and "Usage:" in r.stderr
    r = _cli(str(tmp_path / "missing.txt"))
    assert r.returncode != 0 and "could not" in r.stderr
