"""esthera_amd - MI355X (gfx950) particle-filter engine.

The engine is the C-ABI shared library esthera_amd/libesthera.so (hand-written
HIP kernels, include/esthera.h).  This package is its Python host mirror: a
ctypes binding plus thin classes that read like the reference's driver
(src/partfilt/opencl/pf_generic.c: particle_filter(m, G, t, thr, file)).

There is no fallback: if the library is missing or no GPU is present, the
calls raise.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libesthera.so")

# --------------------------------------------------------------- constants
ESTH_OK = 0
ESTH_ERR_INVALID = -1
ESTH_ERR_BUDGET = -4
ESTH_ERR_UNSUPPORTED = -5
MODEL_ROBOTARM5, MODEL_UNICYCLE, MODEL_DH6, MODEL_UNICYCLE10, MODEL_ROBOTARM5_CAMERA, MODEL_RANDOMWALK3 = range(6)
TOPO_NONE, TOPO_RING, TOPO_TORUS, TOPO_ALL2ALL = 0, 1, 2, 3
RESAMPLE_MULTINOMIAL, RESAMPLE_SYSTEMATIC, RESAMPLE_ALIAS = 0, 1, 2
WEIGHTS_LINEAR, WEIGHTS_LOG = 0, 1
RNG_MTGP_COMPAT, RNG_MTGP_PER_FILTER, RNG_MTGP_PER_FILTER_LONG = 0, 1, 2
LAYOUT_SOA, LAYOUT_AOS, LAYOUT_PACKED = 0, 1, 2
FUSION_AUTO, FUSION_OFF, FUSION_ON = -1, 0, 1
FRAME_FLOATS = 336
MAX_LANDMARKS = 64
STATE_DIM = {MODEL_ROBOTARM5: 9, MODEL_UNICYCLE: 3, MODEL_DH6: 12, MODEL_UNICYCLE10: 3, MODEL_ROBOTARM5_CAMERA: 9,
             MODEL_RANDOMWALK3: 3}
NOISE_DIM = {MODEL_ROBOTARM5: 9, MODEL_UNICYCLE: 3, MODEL_DH6: 6, MODEL_UNICYCLE10: 3, MODEL_ROBOTARM5_CAMERA: 9,
             MODEL_RANDOMWALK3: 3}


# ------------------------------------------------------------- structures
class FilterDesc(C.Structure):
    _fields_ = [("model", C.c_int), ("num_particles", C.c_uint32), ("num_blocks", C.c_uint32),
                ("num_transfer", C.c_uint32), ("topology", C.c_int), ("torus_dimx", C.c_int),
                ("torus_dimy", C.c_int), ("resampler", C.c_int), ("resampling_threshold", C.c_float),
                ("adaptive", C.c_int), ("weights", C.c_int), ("rng", C.c_int), ("seed", C.c_uint32),
                ("carry_weights", C.c_int), ("state_layout", C.c_int)]


class Estimate(C.Structure):
    _fields_ = [("state_dim", C.c_int), ("mean", C.c_double * 16), ("best", C.c_float * 16),
                ("best_weight", C.c_float), ("best_block", C.c_int), ("log_weight_max", C.c_double),
                ("weight_sum", C.c_double), ("step", C.c_uint64)]


class RobotArm5Control(C.Structure):
    _fields_ = [("angles", C.c_float * 5)]


class RobotArm5Measurement(C.Structure):
    _fields_ = [("angles", C.c_float * 5), ("x", C.c_float), ("y", C.c_float)]


class UnicycleParams(C.Structure):
    _fields_ = [("num_landmarks", C.c_int), ("sigma_x", C.c_float), ("sigma_y", C.c_float),
                ("sigma_theta", C.c_float), ("sigma_range", C.c_float), ("sigma_bearing", C.c_float),
                ("lx", C.c_float * MAX_LANDMARKS), ("ly", C.c_float * MAX_LANDMARKS)]


class UnicycleControl(C.Structure):
    _fields_ = [("v", C.c_float), ("omega", C.c_float)]


class UnicycleMeasurement(C.Structure):
    _fields_ = [("range", C.c_float * MAX_LANDMARKS), ("bearing", C.c_float * MAX_LANDMARKS)]


class DH6Params(C.Structure):
    _fields_ = [("a", C.c_float * 6), ("d", C.c_float * 6), ("alpha", C.c_float * 6),
                ("sigma_acc", C.c_float), ("sigma_pos", C.c_float), ("sigma_enc", C.c_float)]


class DH6Control(C.Structure):
    _fields_ = [("unused", C.c_float)]


class DH6Measurement(C.Structure):
    _fields_ = [("p", C.c_float * 3), ("q", C.c_float * 6)]


class Unicycle10Params(C.Structure):
    _fields_ = [("sx", C.c_float * 10), ("sy", C.c_float * 10), ("sz", C.c_float * 10), ("noise_v", C.c_float),
                ("noise_omega", C.c_float), ("noise_gamma", C.c_float)]


class Unicycle10Measurement(C.Structure):
    _fields_ = [("range", C.c_float * 10)]


class CameraParams(C.Structure):
    """Camera intrinsics of robotarm5_camera: focal lengths, principal point,
    distortion kc1, kc2 (radial), kc3, kc4 (tangential)."""
    _fields_ = [("fc", C.c_float * 2), ("cc", C.c_float * 2), ("kc", C.c_float * 4)]


class RandomWalk3Params(C.Structure):
    _fields_ = [("sigma_q", C.c_float), ("sigma_r", C.c_float)]


class RandomWalk3Control(C.Structure):
    _fields_ = [("v", C.c_float * 3)]


class RandomWalk3Measurement(C.Structure):
    _fields_ = [("p", C.c_float * 3)]


class Result(C.Structure):
    _fields_ = [("request_id", C.c_uint64), ("track_id", C.c_uint64), ("status", C.c_int), ("kind", C.c_int),
                ("submit_us", C.c_double), ("done_us", C.c_double), ("estimate", Estimate)]


class ServeWorkload(C.Structure):
    _fields_ = [("rate", C.c_double), ("duration_s", C.c_double), ("open_fraction", C.c_double),
                ("unicycle_fraction", C.c_double), ("min_particles", C.c_uint32), ("max_particles", C.c_uint32),
                ("mean_lifetime", C.c_double), ("initial_tracks", C.c_int), ("landmarks", C.c_int),
                ("topology", C.c_int), ("num_transfer", C.c_int), ("resampler", C.c_int), ("weights", C.c_int),
                ("seed", C.c_uint32), ("plain", C.c_int), ("hbm_bytes", C.c_size_t), ("max_tracks", C.c_int),
                ("max_inflight", C.c_int), ("us_per_mparticle", C.c_void_p), ("ra_controls", C.c_void_p),
                ("ra_measurements", C.c_void_p), ("ra_dt", C.c_void_p), ("ra_steps", C.c_int)]


class ServeReport(C.Structure):
    _fields_ = [("requests", C.c_uint64), ("opens", C.c_uint64), ("updates", C.c_uint64), ("closes", C.c_uint64),
                ("failed", C.c_uint64), ("backlog", C.c_uint64), ("wall_s", C.c_double),
                ("updates_per_s", C.c_double), ("particle_updates_per_s", C.c_double),
                ("mean_particles", C.c_double), ("p50_us", C.c_double), ("p99_us", C.c_double),
                ("mean_us", C.c_double), ("max_us", C.c_double), ("peak_tracks", C.c_int),
                ("hbm_peak", C.c_size_t), ("waves", C.c_uint64), ("mean_wave_updates", C.c_double)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


PARAMS_TYPE = {MODEL_ROBOTARM5: None, MODEL_UNICYCLE: UnicycleParams, MODEL_DH6: DH6Params,
               MODEL_UNICYCLE10: Unicycle10Params, MODEL_ROBOTARM5_CAMERA: CameraParams,
               MODEL_RANDOMWALK3: RandomWalk3Params}
CONTROL_TYPE = {MODEL_ROBOTARM5: RobotArm5Control, MODEL_UNICYCLE: UnicycleControl, MODEL_DH6: DH6Control,
                MODEL_UNICYCLE10: UnicycleControl, MODEL_ROBOTARM5_CAMERA: RobotArm5Control,
                MODEL_RANDOMWALK3: RandomWalk3Control}
MEAS_TYPE = {MODEL_ROBOTARM5: RobotArm5Measurement, MODEL_UNICYCLE: UnicycleMeasurement,
             MODEL_DH6: DH6Measurement, MODEL_UNICYCLE10: Unicycle10Measurement,
             MODEL_ROBOTARM5_CAMERA: RobotArm5Measurement, MODEL_RANDOMWALK3: RandomWalk3Measurement}


class EsthError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"esthera error {code}: {msg}")
        self.code = code


_lib = None

_P = C.c_void_p
_SIGS = {
    "esth_last_error": ([], C.c_char_p),
    "esth_version": ([], C.c_int),
    "esth_model_lookup": ([C.c_char_p], C.c_int),
    "esth_model_info": ([C.c_int, _P, _P, _P, _P, _P], C.c_int),
    "esth_model_build_frame": ([C.c_int, _P, _P, _P, C.c_float, _P], C.c_int),
    "esth_filter_desc_default": ([_P, C.c_int], None),
    "esth_engine_create": ([C.c_int, C.c_size_t, _P], C.c_int),
    "esth_engine_destroy": ([_P], C.c_int),
    "esth_engine_memory": ([_P, _P, _P], C.c_int),
    "esth_engine_set_batch_fusion": ([_P, C.c_uint32, C.c_int], C.c_int),
    "esth_engine_batch_fusion": ([_P, C.c_uint32, _P], C.c_int),
    "esth_engine_set_batch_resample_lanes": ([_P, C.c_uint32, C.c_int], C.c_int),
    "esth_engine_batch_resample_lanes": ([_P, C.c_uint32, _P], C.c_int),
    "esth_filter_create": ([_P, _P, _P, C.c_size_t, _P], C.c_int),
    "esth_filter_destroy": ([_P], C.c_int),
    "esth_filter_init_uniform": ([_P, _P], C.c_int),
    "esth_filter_init": ([_P, _P], C.c_int),
    "esth_filter_step": ([_P, _P, _P, C.c_float], C.c_int),
    "esth_filter_sync": ([_P], C.c_int),
    "esth_filter_estimate": ([_P, _P], C.c_int),
    "esth_filter_read_particles": ([_P, _P, _P], C.c_int),
    "esth_filter_set_timing": ([_P, C.c_int], C.c_int),
    "esth_filter_set_timing_every": ([_P, C.c_int], C.c_int),
    "esth_filter_stage_times": ([_P, _P], C.c_int),
    "esth_filter_snapshot_size": ([_P, _P], C.c_int),
    "esth_filter_snapshot": ([_P, _P, C.c_size_t], C.c_int),
    "esth_filter_restore": ([_P, _P, C.c_size_t], C.c_int),
    "esth_filter_stream": ([_P, _P], C.c_int),
    "esth_filter_use_graph": ([_P, C.c_int], C.c_int),
    "esth_filter_set_prefetch": ([_P, C.c_int], C.c_int),
    "esth_filter_set_fusion": ([_P, C.c_int, _P], C.c_int),
    "esth_particle_filter": ([C.c_uint32, C.c_uint32, C.c_uint32, C.c_float, C.c_char_p, C.c_uint32, C.c_int,
                              _P, _P], C.c_int),
    "esth_trace_load_robotarm": ([C.c_char_p, C.c_int, _P, _P, _P, _P, _P], C.c_int),
    "esth_robotarm5_estimate_error": ([_P, _P], C.c_float),
    "esth_trace_load_unicycle10": ([C.c_char_p, C.c_int, _P, _P, _P, _P, _P], C.c_int),
    "esth_trace_write_robotarm": ([C.c_char_p, C.c_int, _P, _P, _P, _P], C.c_int),
    "esth_trace_write_unicycle10": ([C.c_char_p, C.c_int, _P, _P, _P, _P], C.c_int),
    "esth_unicycle10_filter": ([C.c_uint32, C.c_uint32, C.c_uint32, C.c_float, C.c_char_p, C.c_uint32, C.c_int,
                                _P, _P], C.c_int),
    "esth_unicycle10_params_default": ([_P], None),
    "esth_camera_params_default": ([_P], None),
    "esth_unicycle_scenario": ([C.c_uint32, C.c_int, C.c_float, C.c_int, C.c_float, _P, _P, _P, _P], C.c_int),
    "esth_dh6_scenario": ([C.c_uint32, C.c_int, C.c_float, _P, _P, _P], C.c_int),
    "esth_robotarm5_camera_scenario": ([C.c_uint32, C.c_int, C.c_float, _P, _P, _P, _P], C.c_int),
    "esth_batch_create": ([_P, C.c_int, _P, _P, C.c_size_t, _P], C.c_int),
    "esth_batch_destroy": ([_P], C.c_int),
    "esth_batch_init_uniform": ([_P, _P], C.c_int),
    "esth_batch_step": ([_P, _P, C.c_size_t, _P, C.c_size_t, _P], C.c_int),
    "esth_batch_sync": ([_P], C.c_int),
    "esth_batch_estimate": ([_P, C.c_int, _P], C.c_int),
    "esth_batch_read_particles": ([_P, C.c_int, _P, _P], C.c_int),
    "esth_batch_use_graph": ([_P, C.c_int], C.c_int),
    "esth_batch_step_time": ([_P, _P], C.c_int),
    "esth_batch_step_times": ([_P, _P, C.c_int, _P], C.c_int),
    "esth_batch_info": ([_P, _P, _P, _P], C.c_int),
    "esth_autotune": ([_P, _P, _P, C.c_size_t, _P, _P, C.c_float, _P], C.c_int),
    "esth_autotune_pick": ([_P, C.c_uint32, C.c_uint32, _P], C.c_int),
    "esth_autotune_layout": ([_P, _P, _P, C.c_size_t, _P, _P, C.c_float, _P, _P], C.c_int),
    "esth_engine_time_hook": ([_P, C.c_int, _P, C.c_uint64, C.c_int, _P], C.c_int),
    "esth_state_convert": ([C.c_int, C.c_size_t, C.c_int, _P, C.c_int, _P], C.c_int),
    "esth_server_create": ([_P, C.c_int, C.c_size_t, _P], C.c_int),
    "esth_server_configure": ([_P, C.c_int, C.c_int], C.c_int),
    "esth_server_wave_stats": ([_P, _P, _P, _P], C.c_int),
    "esth_server_read_particles": ([_P, C.c_uint64, _P, _P], C.c_int),
    "esth_serve_simulate": ([_P, _P, _P], C.c_int),
    "esth_server_destroy": ([_P], C.c_int),
    "esth_server_open": ([_P, _P, _P, C.c_size_t, _P, _P, _P], C.c_int),
    "esth_server_update": ([_P, C.c_uint64, _P, _P, C.c_float, _P], C.c_int),
    "esth_server_close": ([_P, C.c_uint64, _P], C.c_int),
    "esth_server_pump": ([_P], C.c_int),
    "esth_server_poll": ([_P, _P, C.c_int, _P], C.c_int),
    "esth_server_drain": ([_P], C.c_int),
    "esth_server_stats": ([_P, _P, _P, _P], C.c_int),
    "esth_set_device": ([C.c_int], C.c_int),
    "esth_dev_alloc": ([C.c_size_t, _P], C.c_int),
    "esth_dev_free": ([_P], C.c_int),
    "esth_memcpy_htod": ([_P, _P, C.c_size_t], C.c_int),
    "esth_memcpy_dtoh": ([_P, _P, C.c_size_t], C.c_int),
    "esth_memset_d": ([_P, C.c_int, C.c_size_t], C.c_int),
    "esth_dev_sync": ([], C.c_int),
    "esth_mtgp_params": ([], C.POINTER(C.c_uint32)),
    "esth_mtgp_seed_state": ([C.c_int, C.c_uint32, _P], C.c_int),
    "esth_stage_mtgp": ([_P, C.c_int, C.c_int, C.c_int, _P, C.c_int], C.c_int),
    "esth_stage_mtgp_steps": ([_P, _P, C.c_int, C.c_int, C.c_int, C.c_int, _P, C.c_int], C.c_int),
    "esth_stage_boxmuller": ([_P, C.c_int], C.c_int),
    "esth_stage_sample_weight": ([C.c_int, _P, C.c_uint32, _P, _P, C.c_int, _P], C.c_int),
    "esth_stage_sort": ([_P, _P, _P, C.c_int, C.c_uint32, C.c_uint32, _P, _P], C.c_int),
    "esth_filter_set_cross_step": ([_P, C.c_int, _P], C.c_int),
    "esth_filter_rng_mode": ([_P, _P, _P], C.c_int),
    "esth_stage_sort_noalloc": ([_P, _P, _P, C.c_int, C.c_uint32, C.c_uint32, _P, _P, _P], C.c_int),
    "esth_stage_argmax": ([_P, C.c_uint32, C.c_int, _P], C.c_int),
    "esth_stage_exchange": ([_P, _P, C.c_int, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int, C.c_int, C.c_int],
                            C.c_int),
    "esth_stage_resample": ([C.c_int, _P, _P, C.c_int, _P, _P, C.c_uint32, C.c_uint32, C.c_float, C.c_int,
                             C.c_int, _P], C.c_int),
    "esth_stage_alias_table": ([_P, C.c_uint32, C.c_uint32, C.c_int, _P, _P], C.c_int),
    "esth_stage_math": ([C.c_int, _P, _P, _P, C.c_int], C.c_int),
    "esth_stage_copy": ([_P, _P, C.c_size_t], C.c_int),
}


def lib():
    """The loaded engine library (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"engine library not built: {LIB_PATH} (python -m esthera_amd.build)")
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in _SIGS.items():
            fn = getattr(L, name, None)
            if fn is None:
                continue
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def check(rc):
    if rc != ESTH_OK:
        raise EsthError(rc, lib().esth_last_error().decode())
    return rc


def ptr(x):
    """void* of a ctypes object, numpy array or DeviceBuffer (None -> NULL)."""
    if x is None:
        return None
    if isinstance(x, DeviceBuffer):
        return x.ptr
    if isinstance(x, np.ndarray):
        return x.ctypes.data_as(C.c_void_p)
    if isinstance(x, int):
        return C.c_void_p(x)
    return C.cast(C.pointer(x), C.c_void_p)


# ------------------------------------------------------------- device memory
class DeviceBuffer:
    """A device allocation made through the engine's C ABI (hipMalloc)."""

    def __init__(self, nbytes):
        p = C.c_void_p()
        check(lib().esth_dev_alloc(nbytes, C.byref(p)))
        self.ptr = p
        self.nbytes = nbytes

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a)
        b = cls(max(a.nbytes, 4))
        if a.nbytes:
            check(lib().esth_memcpy_htod(b.ptr, a.ctypes.data_as(C.c_void_p), a.nbytes))
        return b

    def to_array(self, dtype, count):
        out = np.empty(count, dtype=dtype)
        if out.nbytes:
            check(lib().esth_memcpy_dtoh(out.ctypes.data_as(C.c_void_p), self.ptr, out.nbytes))
        return out

    def upload(self, a):
        a = np.ascontiguousarray(a)
        check(lib().esth_memcpy_htod(self.ptr, a.ctypes.data_as(C.c_void_p), a.nbytes))

    def free(self):
        if self.ptr:
            lib().esth_dev_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def device_sync():
    check(lib().esth_dev_sync())


# ------------------------------------------------------------------ engine
class Engine:
    """One engine per GPU; owns the HBM budget."""

    def __init__(self, device=0, hbm_budget=0):
        h = C.c_void_p()
        check(lib().esth_engine_create(device, hbm_budget, C.byref(h)))
        self.h = h

    def memory(self):
        used, budget = C.c_size_t(), C.c_size_t()
        check(lib().esth_engine_memory(self.h, C.byref(used), C.byref(budget)))
        return used.value, budget.value

    def set_batch_fusion(self, m, mode):
        """Batches/servers: sub-filter size class m sampled and sorted in one
        launch (FUSION_ON), two (FUSION_OFF) or as the engine decides (FUSION_AUTO)."""
        check(lib().esth_engine_set_batch_fusion(self.h, m, mode))

    def batch_fusion(self, m):
        mode = C.c_int()
        check(lib().esth_engine_batch_fusion(self.h, m, C.byref(mode)))
        return mode.value

    def set_batch_resample_lanes(self, m, lanes):
        """Batches/servers: lanes per sub-filter of size class m's resampling:
        32 (m = 64), 64 (one wave), 128 (m >= 256); 0 the default."""
        check(lib().esth_engine_set_batch_resample_lanes(self.h, m, lanes))

    def batch_resample_lanes(self, m):
        lanes = C.c_int()
        check(lib().esth_engine_batch_resample_lanes(self.h, m, C.byref(lanes)))
        return lanes.value

    def close(self):
        if self.h:
            lib().esth_engine_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def filter_desc(model, m=256, G=16, t=1, topology=TOPO_RING, resampler=RESAMPLE_MULTINOMIAL, threshold=1.0,
                adaptive=0, weights=WEIGHTS_LINEAR, rng=RNG_MTGP_COMPAT, seed=1234, dimx=0, dimy=0, carry=0,
                layout=LAYOUT_SOA):
    d = FilterDesc()
    d.model, d.num_particles, d.num_blocks, d.num_transfer = model, m, G, t
    d.topology, d.torus_dimx, d.torus_dimy = topology, dimx, dimy
    d.resampler, d.resampling_threshold, d.adaptive = resampler, threshold, adaptive
    d.weights, d.rng, d.seed, d.carry_weights = weights, rng, seed, carry
    d.state_layout = layout
    return d


class Filter:
    """An island-model filter: m particles per sub-filter, G sub-filters.

    Mirrors particle_filter() of pf_generic.c:245-887: create, set the
    initial particles, then step(control, measurement, dt) per trace row."""

    def __init__(self, engine, desc, params=None):
        self.engine = engine
        self.desc = desc
        self.S = STATE_DIM[desc.model]
        self.N = desc.num_particles * desc.num_blocks
        self._params = params
        h = C.c_void_p()
        psize = C.sizeof(params) if params is not None else 0
        check(lib().esth_filter_create(engine.h, ptr(desc), ptr(params), psize, C.byref(h)))
        self.h = h

    def init_uniform(self, state):
        s = np.ascontiguousarray(state, dtype=np.float32)
        check(lib().esth_filter_init_uniform(self.h, ptr(s)))

    def init(self, soa):
        s = np.ascontiguousarray(soa, dtype=np.float32)
        assert s.size == self.S * self.N
        check(lib().esth_filter_init(self.h, ptr(s)))

    def step(self, control, measurement, dt):
        check(lib().esth_filter_step(self.h, ptr(control), ptr(measurement), dt))

    def sync(self):
        check(lib().esth_filter_sync(self.h))

    def estimate(self):
        e = Estimate()
        check(lib().esth_filter_estimate(self.h, C.byref(e)))
        return e

    def particles(self):
        st = np.empty(self.S * self.N, np.float32)
        w = np.empty(self.N, np.float32)
        check(lib().esth_filter_read_particles(self.h, ptr(st), ptr(w)))
        return st, w

    def set_timing(self, level=1, every=1):
        """0 off, 1 time the sample+weight kernel, 2 time every stage; one
        step in `every` is timed."""
        check(lib().esth_filter_set_timing(self.h, int(level)))
        check(lib().esth_filter_set_timing_every(self.h, int(every)))

    def stage_times(self):
        us = np.zeros(6, np.float32)
        check(lib().esth_filter_stage_times(self.h, ptr(us)))
        return us

    def set_prefetch(self, on=True):
        check(lib().esth_filter_set_prefetch(self.h, 1 if on else 0))

    def set_cross_step(self, on=True):
        """One launch per step across the step boundary where it applies;
        returns whether it is taken."""
        a = C.c_int()
        check(lib().esth_filter_set_cross_step(self.h, 1 if on else 0, C.byref(a)))
        return bool(a.value)

    def rng_mode(self):
        """(inline, block_steps): whether the next step's MTGP32 words come
        from the sampling launch, else the block length of the second stream."""
        a, b = C.c_int(), C.c_int()
        check(lib().esth_filter_rng_mode(self.h, C.byref(a), C.byref(b)))
        return bool(a.value), b.value

    def set_fusion(self, on=True):
        """Sample+weight and sort in one launch where it pays (default);
        returns whether steps use it."""
        active = C.c_int(0)
        check(lib().esth_filter_set_fusion(self.h, 1 if on else 0, C.byref(active)))
        return bool(active.value)

    def use_graph(self, on=True):
        check(lib().esth_filter_use_graph(self.h, 1 if on else 0))

    def snapshot(self):
        n = C.c_size_t()
        check(lib().esth_filter_snapshot_size(self.h, C.byref(n)))
        buf = np.empty(n.value, np.uint8)
        check(lib().esth_filter_snapshot(self.h, ptr(buf), n.value))
        return buf

    def restore(self, buf):
        check(lib().esth_filter_restore(self.h, ptr(buf), buf.nbytes))

    def close(self):
        if self.h:
            lib().esth_filter_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many filters stepped together (one graph launch per step)."""

    def __init__(self, engine, descs, params=None):
        self.engine = engine
        self.n = len(descs)
        arr = (FilterDesc * self.n)(*descs)
        self.descs = list(descs)
        self.S = STATE_DIM[descs[0].model]
        self._params = params
        h = C.c_void_p()
        if params is not None:
            parr = (type(params[0]) * self.n)(*params)
            self._parr = parr
            check(lib().esth_batch_create(engine.h, self.n, arr, C.cast(parr, C.c_void_p), C.sizeof(type(params[0])),
                                          C.byref(h)))
        else:
            check(lib().esth_batch_create(engine.h, self.n, arr, None, 0, C.byref(h)))
        self.h = h

    def init_uniform(self, states):
        s = np.ascontiguousarray(states, dtype=np.float32).reshape(self.n, self.S)
        check(lib().esth_batch_init_uniform(self.h, ptr(s)))

    def step(self, controls, measurements, dts):
        """controls / measurements: ctypes arrays of the model structs (one per filter)."""
        d = np.ascontiguousarray(dts, dtype=np.float32)
        cptr = None if controls is None else C.cast(controls, C.c_void_p)
        cs = 0 if controls is None else C.sizeof(controls._type_)
        check(lib().esth_batch_step(self.h, cptr, cs, C.cast(measurements, C.c_void_p),
                                    C.sizeof(measurements._type_), ptr(d)))

    def sync(self):
        check(lib().esth_batch_sync(self.h))

    def use_graph(self, on=True):
        check(lib().esth_batch_use_graph(self.h, 1 if on else 0))

    def estimate(self, i):
        e = Estimate()
        check(lib().esth_batch_estimate(self.h, i, C.byref(e)))
        return e

    def particles(self, i):
        N = self.descs[i].num_particles * self.descs[i].num_blocks
        st = np.empty(self.S * N, np.float32)
        w = np.empty(N, np.float32)
        check(lib().esth_batch_read_particles(self.h, i, ptr(st), ptr(w)))
        return st, w

    def step_times(self, max_steps=8192):
        us = np.zeros(max_steps, np.float32)
        n = C.c_int()
        check(lib().esth_batch_step_times(self.h, ptr(us), max_steps, C.byref(n)))
        return us[: n.value]

    def info(self):
        p, s, g = C.c_uint64(), C.c_int(), C.c_uint32()
        check(lib().esth_batch_info(self.h, C.byref(p), C.byref(s), C.byref(g)))
        return p.value, s.value, g.value

    def close(self):
        if self.h:
            lib().esth_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching request loop over resident tracks."""

    def __init__(self, engine, max_tracks, hbm_bytes, max_inflight=2, max_wave_tracks=1 << 30):
        self.engine = engine
        h = C.c_void_p()
        check(lib().esth_server_create(engine.h, max_tracks, hbm_bytes, C.byref(h)))
        self.h = h
        check(lib().esth_server_configure(self.h, max_inflight, max_wave_tracks))
        self.S = {}

    def open(self, desc, params, init_state):
        """Queue a track; returns (track_id, request_id)."""
        s = np.zeros(16, np.float32)
        init = np.asarray(init_state, np.float32).ravel()
        s[: init.size] = init
        tid, rid = C.c_uint64(), C.c_uint64()
        psize = C.sizeof(params) if params is not None else 0
        check(lib().esth_server_open(self.h, ptr(desc), ptr(params), psize, ptr(s), C.byref(tid), C.byref(rid)))
        self.S[tid.value] = (STATE_DIM[desc.model], desc.num_particles * desc.num_blocks)
        return tid.value, rid.value

    def update(self, track, control, measurement, dt):
        rid = C.c_uint64()
        check(lib().esth_server_update(self.h, track, ptr(control), ptr(measurement), dt, C.byref(rid)))
        return rid.value

    def close_track(self, track):
        rid = C.c_uint64()
        check(lib().esth_server_close(self.h, track, C.byref(rid)))
        return rid.value

    def pump(self):
        check(lib().esth_server_pump(self.h))

    def drain(self):
        check(lib().esth_server_drain(self.h))

    def poll(self, max_results=4096):
        out = (Result * max_results)()
        n = C.c_int()
        check(lib().esth_server_poll(self.h, out, max_results, C.byref(n)))
        return [out[i] for i in range(n.value)]

    def stats(self):
        used, open_tracks, queued = C.c_size_t(), C.c_int(), C.c_int()
        check(lib().esth_server_stats(self.h, C.byref(used), C.byref(open_tracks), C.byref(queued)))
        return used.value, open_tracks.value, queued.value

    def wave_stats(self):
        waves, upd, peak = C.c_uint64(), C.c_uint64(), C.c_size_t()
        check(lib().esth_server_wave_stats(self.h, C.byref(waves), C.byref(upd), C.byref(peak)))
        return waves.value, upd.value, peak.value

    def particles(self, track):
        S, N = self.S[track]
        st = np.empty(S * N, np.float32)
        w = np.empty(N, np.float32)
        check(lib().esth_server_read_particles(self.h, track, ptr(st), ptr(w)))
        return st, w

    def close(self):
        if self.h:
            lib().esth_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def serve_simulate(engine, workload, ra_trace=None, us_per_mparticle=None):
    """Drive a seeded request stream (esth_serve_simulate); ra_trace =
    (controls, measurements, dts) from load_robotarm_trace for robot-arm tracks."""
    keep = []
    if us_per_mparticle is not None:
        us = np.ascontiguousarray(us_per_mparticle, dtype=np.float32)
        keep.append(us)
        workload.us_per_mparticle = us.ctypes.data
    if ra_trace is not None:
        meas, ctrl, dts = ra_trace
        carr = (RobotArm5Control * len(ctrl))(*ctrl)
        marr = (RobotArm5Measurement * len(meas))(*meas)
        d = np.ascontiguousarray(dts, dtype=np.float32)
        keep += [carr, marr, d]
        workload.ra_controls = C.cast(carr, C.c_void_p)
        workload.ra_measurements = C.cast(marr, C.c_void_p)
        workload.ra_dt = d.ctypes.data
        workload.ra_steps = len(ctrl)
    rep_ = ServeReport()
    check(lib().esth_serve_simulate(engine.h, C.byref(workload), C.byref(rep_)))
    return rep_.as_dict()


def autotune(engine, desc, params, control, measurement, dt):
    us = np.zeros(5, np.float32)
    psize = C.sizeof(params) if params is not None else 0
    check(lib().esth_autotune(engine.h, ptr(desc), ptr(params), psize, ptr(control), ptr(measurement), dt, ptr(us)))
    return us


def autotune_layout(engine, desc, params, control, measurement, dt):
    """Device time per step (us) of desc's filter in layouts SOA, AOS, PACKED, and the fastest."""
    us = np.zeros(3, np.float32)
    best = C.c_int()
    psize = C.sizeof(params) if params is not None else 0
    check(lib().esth_autotune_layout(engine.h, ptr(desc), ptr(params), psize, ptr(control), ptr(measurement), dt,
                                     ptr(us), C.byref(best)))
    return us, best.value


def time_hook(engine, model, frame, n, reps=20):
    """Device time (us) of the model's sample+weight arithmetic alone over n
    particles (no memory access but one store per pair): the floor of the
    sample+weight launches' arithmetic on this device."""
    fr = np.ascontiguousarray(frame, dtype=np.float32)
    if fr.size != FRAME_FLOATS:
        raise ValueError(f"frame must hold {FRAME_FLOATS} floats")
    us = C.c_float()
    check(lib().esth_engine_time_hook(engine.h, model, ptr(fr), n, reps, C.byref(us)))
    return us.value


def state_convert(state, state_dim, n, src_layout, dst_layout):
    """A host buffer of n particles of state_dim floats from one LAYOUT_* to another."""
    src = np.ascontiguousarray(state, dtype=np.float32).reshape(-1)
    assert src.size == state_dim * n
    dst = np.empty_like(src)
    check(lib().esth_state_convert(state_dim, n, src_layout, ptr(src), dst_layout, ptr(dst)))
    return dst


def autotune_pick(us_per_mparticle, particles, min_m=256):
    us = np.ascontiguousarray(us_per_mparticle, dtype=np.float32)
    m = C.c_uint32()
    check(lib().esth_autotune_pick(ptr(us), particles, min_m, C.byref(m)))
    return m.value


def particle_filter(m, G, t, threshold, trace_path, seed=1234, resampler=RESAMPLE_MULTINOMIAL):
    """particle_filter(m, G, t, thr, file) of pf_generic.c:245 for the robot
    arm: returns (mean estimate error over steps, steps)."""
    err = C.c_double()
    steps = C.c_int()
    check(lib().esth_particle_filter(m, G, t, threshold, trace_path.encode(), seed, resampler, C.byref(err),
                                     C.byref(steps)))
    return err.value, steps.value


def unicycle_scenario(seed, num_landmarks, box, steps, dt=0.1):
    p = UnicycleParams()
    ctrl = (UnicycleControl * steps)()
    meas = (UnicycleMeasurement * steps)()
    truth = np.zeros(steps * 3, np.float32)
    check(lib().esth_unicycle_scenario(seed, num_landmarks, box, steps, dt, ptr(p), ctrl, meas, ptr(truth)))
    return p, ctrl, meas, truth.reshape(steps, 3)


def dh6_scenario(seed, steps, dt=0.01):
    p = DH6Params()
    meas = (DH6Measurement * steps)()
    truth = np.zeros((steps + 1) * 12, np.float32)
    check(lib().esth_dh6_scenario(seed, steps, dt, ptr(p), meas, ptr(truth)))
    return p, meas, truth.reshape(steps + 1, 12)


def randomwalk3_scenario(seed, steps, dt=0.1, sigma_q=0.05, sigma_r=0.2):
    """Seeded random-walk trace: (params, controls, measurements, truth rows
    steps x 3); the truth walks with a constant velocity plus N(0, sigma_q)
    steps, each measurement is the truth plus N(0, sigma_r) noise."""
    rng = np.random.default_rng(seed)
    p = RandomWalk3Params(sigma_q, sigma_r)
    v = rng.uniform(-1.0, 1.0, 3).astype(np.float32)
    ctrl = (RandomWalk3Control * steps)()
    meas = (RandomWalk3Measurement * steps)()
    truth = np.zeros((steps, 3), np.float32)
    x = np.zeros(3, np.float64)
    for k in range(steps):
        x = x + v * dt + rng.normal(0.0, sigma_q, 3)
        truth[k] = x
        ctrl[k].v[:] = v.tolist()
        meas[k].p[:] = (x + rng.normal(0.0, sigma_r, 3)).astype(np.float32).tolist()
    return p, ctrl, meas, truth


def robotarm5_camera_scenario(seed, steps, dt=0.1, params=None):
    """Seeded camera-observed arm trace: (controls, measurements, truth rows (steps+1) x 9)."""
    ctrl = (RobotArm5Control * steps)()
    meas = (RobotArm5Measurement * steps)()
    truth = np.zeros((steps + 1) * 9, np.float32)
    check(lib().esth_robotarm5_camera_scenario(seed, steps, dt, ptr(params), ctrl, meas, ptr(truth)))
    return ctrl, meas, truth.reshape(steps + 1, 9)


def load_robotarm_trace(path, max_rows=100000):
    meas = (RobotArm5Measurement * max_rows)()
    ctrl = (RobotArm5Control * max_rows)()
    dt = np.zeros(max_rows, np.float32)
    truth = np.zeros(max_rows * 9, np.float32)
    rows = C.c_int()
    check(lib().esth_trace_load_robotarm(path.encode(), max_rows, meas, ctrl, ptr(dt), ptr(truth), C.byref(rows)))
    n = rows.value
    return meas[:n], ctrl[:n], dt[:n], truth[: n * 9].reshape(n, 9)


def write_robotarm_trace(path, meas, ctrl, dt, truth):
    """Rows in the robot-arm trace format (truth: rows x 9, vx and vy not written)."""
    n = len(meas)
    m = (RobotArm5Measurement * n)(*meas)
    c = (RobotArm5Control * n)(*ctrl)
    d = np.ascontiguousarray(dt, dtype=np.float32)
    t = np.ascontiguousarray(truth, dtype=np.float32).reshape(n, 9)
    check(lib().esth_trace_write_robotarm(path.encode(), n, m, c, ptr(d), ptr(t)))


def write_unicycle10_trace(path, meas, ctrl, dt, truth):
    n = len(meas)
    m = (Unicycle10Measurement * n)(*meas)
    c = (UnicycleControl * n)(*ctrl)
    d = np.ascontiguousarray(dt, dtype=np.float32)
    t = np.ascontiguousarray(truth, dtype=np.float32).reshape(n, 3)
    check(lib().esth_trace_write_unicycle10(path.encode(), n, m, c, ptr(d), ptr(t)))


CLI_PATH = os.path.join(HERE, "esthera_pf")


def load_unicycle10_trace(path, max_rows=100000):
    """The reference unicycle's 16-column input: (measurements, controls, dt, truth rows x 3)."""
    meas = (Unicycle10Measurement * max_rows)()
    ctrl = (UnicycleControl * max_rows)()
    dt = np.zeros(max_rows, np.float32)
    truth = np.zeros(max_rows * 3, np.float32)
    rows = C.c_int()
    check(lib().esth_trace_load_unicycle10(path.encode(), max_rows, meas, ctrl, ptr(dt), ptr(truth), C.byref(rows)))
    n = rows.value
    return meas[:n], ctrl[:n], dt[:n], truth[: n * 3].reshape(n, 3)


def unicycle10_params_default():
    p = Unicycle10Params()
    lib().esth_unicycle10_params_default(C.byref(p))
    return p


def camera_params_default():
    """The reference's camera calibration (pf_model_soa_packed_code.cuh:33-43)."""
    p = CameraParams()
    lib().esth_camera_params_default(C.byref(p))
    return p


def unicycle10_filter(m, G, t, threshold, trace_path, seed=1234, resampler=RESAMPLE_MULTINOMIAL):
    """The reference unicycle program's loop on the engine: (mean best-particle
    distance to the truth, steps)."""
    err = C.c_double()
    steps = C.c_int()
    check(lib().esth_unicycle10_filter(m, G, t, threshold, trace_path.encode(), seed, resampler, C.byref(err),
                                       C.byref(steps)))
    return err.value, steps.value


def build_frame(model, params, control, measurement, dt):
    fr = np.zeros(FRAME_FLOATS, np.float32)
    check(lib().esth_model_build_frame(model, ptr(params), ptr(control), ptr(measurement), dt, ptr(fr)))
    return fr


def mtgp_params():
    p = lib().esth_mtgp_params()
    return np.ctypeslib.as_array(p, shape=(200 * 52,)).reshape(200, 52).copy()


def mtgp_seed_state(set_idx, seed):
    st = np.zeros(351, np.uint32)
    check(lib().esth_mtgp_seed_state(set_idx, seed & 0xFFFFFFFF, ptr(st)))
    return st
